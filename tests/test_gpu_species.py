"""GPU parity of the multi-species flow path (N_SPEC = 9, real gas, Chemkin
fits with Wilke mixing, BC_JET) and of its coupling to the chemistry.

Bar: <= 1e-12 relative (to each field's largest magnitude) per conserved
variable, primitive and partial density after one RK3 step and after an
NS_solver step with chemistry.  Differences come only from the device
exp/log/sqrt used by the transport fits, the Wilke weights and the
chemistry (a few ulps per call); the fraction of bit-identical values is
reported.
"""
import numpy as np
import pytest

from conftest import load_package, gpu_available
import flowcase as fc

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def pkg():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_package()


def pair(pkg, case):
    import oracle
    cfg = case[0]
    g = pkg.FlowSolver(cfg)
    o = oracle.Oracle(pkg, cfg)
    fc.setup_jet(g, case)
    fc.setup_jet(o, case)
    return g, o


def compare_all(g, o, names=("F", "D", "U", "V", "W", "T", "P", "ET", "DI"), tol=TOL):
    rep = {}
    for name in names:
        a, b = g.download(name), o.download(name)
        if name in fc.PRIMS or name == "DI":
            a, b = fc.interior(a), fc.interior(b)
        if a.ndim == 4:
            for m in range(a.shape[3]):
                rep[f"{name}{m + 1}"] = fc.compare(a[..., m], b[..., m], name)
        else:
            rep[name] = fc.compare(a, b, name)
    print({k: (f"{v[0]:.1e}", f"{v[1]:.3f}") for k, v in rep.items()})
    bad = {k: v for k, v in rep.items() if not (v[0] <= tol)}
    assert not bad, f"parity failure (max rel diff, bit-identical fraction): {bad}"
    return rep


def test_setup_state(pkg):
    g, o = pair(pkg, fc.small_jet())
    compare_all(g, o)
    g.close()


def test_transport_coefficients(pkg):
    g, o = pair(pkg, fc.small_jet())
    g.comput_Amu_Amk_AMD()
    o.comput_Amu_Amk_AMD()
    compare_all(g, o, names=("AMU", "AMK", "AMD"))
    g.close()


def test_species_right_hand_sides(pkg):
    g, o = pair(pkg, fc.small_jet())
    g.comput_Amu_Amk_AMD()
    o.comput_Amu_Amk_AMD()
    o.exchange_all()
    g.du_species()
    dus, qp = o.du_species()
    for name, a, b in (("DUS", g.download("DUS"), dus), ("QP", g.download("QP"), qp)):
        for m in range(a.shape[3]):
            d, same = fc.compare(a[..., m], b[..., m], name)
            print(name, m, d, same)
            assert d <= TOL, (name, m, d)
    g.close()


@pytest.mark.parametrize("iflag", [0, 1])
def test_rk3_step(pkg, iflag):
    case = fc.small_jet(iflag=iflag)
    g, o = pair(pkg, case)
    dt = case[5]["dt"]
    g.flow_time_advance(dt)
    o.flow_time_advance(dt)
    compare_all(g, o)
    assert g.negative_et_count() == o.negative_et_count()
    g.close()


def test_rk3_step_wide_rows(pkg):
    """Rows of 160 cells, so that whole wavefronts have the central xi
    stencil (i = 65..128 on even rows, 33..96 on odd ones): the central xi,
    eta and zeta forms of the multi-species viscous kernels and of the flow's
    viscous kernels run beside the face forms; one RK3 step against the
    oracle."""
    case = fc.small_jet(nx=160, ny=12, nz=10)
    g, o = pair(pkg, case)
    dt = case[5]["dt"]
    g.flow_time_advance(dt)
    o.flow_time_advance(dt)
    compare_all(g, o)
    assert g.negative_et_count() == o.negative_et_count()
    g.close()


def test_column_planes_bitwise(pkg):
    """The k-marching viscous-flux and update kernels (non-periodic z, where
    the windows meet the one-sided closures) give the one-thread-per-cell
    launch bit for bit for chunk lengths that do and do not divide nz."""
    case = fc.small_jet(nz=22)
    dt = case[5]["dt"]
    ref = None
    for kc in (0, 1, 5, 16, 22):
        g = pkg.FlowSolver(case[0])
        g.set_option(pkg.OPT_COLUMN_PLANES, kc)
        fc.setup_jet(g, case)
        g.flow_time_advance(dt)
        snap = {n: g.download(n) for n in ("F", "DI", "T", "P")}
        g.close()
        if ref is None:
            ref = snap
            continue
        for n in snap:
            assert np.array_equal(snap[n].view(np.uint64), ref[n].view(np.uint64)), (kc, n)


def test_stage_pieces_equal_fused_step(pkg):
    case = fc.small_jet()
    dt = case[5]["dt"]
    a = pkg.FlowSolver(case[0])          # one library context at a time
    fc.setup_jet(a, case)
    a.flow_time_advance(dt)
    fa = {n: a.download(n) for n in ("F", "DI", "T", "P")}
    a.close()
    b = pkg.FlowSolver(case[0])
    fc.setup_jet(b, case)
    b.save_stage0()
    for krk in (1, 2, 3):
        b.stage_begin(krk)
        b.Scalar_time_adv_RK3(krk, dt)
        b.du_inviscous()
        b.du_viscous()
        b.OCFD_time_adv_RK3(krk, dt)
        b.comput_flow_variables()
        b.OCFD_bc()
    for n, v in fa.items():
        assert np.array_equal(v, b.download(n)), n
    b.close()


def test_perfect_gas_sutherland_mixture(pkg):
    """N_SPEC = 9 with the perfect-gas thermo and Sutherland/Pr/Sc transport."""
    case = list(fc.small_jet())
    cfg = case[0]
    cfg.if_real_gas = 0
    cfg.flag_vis = cfg.flag_cond = cfg.flag_diff = 0
    g, o = pair(pkg, tuple(case))
    dt = case[5]["dt"]
    for _ in range(2):
        g.flow_time_advance(dt)
        o.flow_time_advance(dt)
    compare_all(g, o)
    g.close()


def test_concurrent_streams_bitwise(pkg):
    """The species chain on a second stream beside the flow sweeps
    (OPT_CONCURRENT = 1, the default) gives the one-stream run bit for bit,
    eagerly and under graph replay, over two RK3 steps of the jet."""
    case = fc.small_jet()
    dt = case[5]["dt"]
    out = []
    for conc, graph in ((0, 0), (1, 0), (1, 1)):
        g = pkg.FlowSolver(case[0])
        g.set_option(pkg.OPT_CONCURRENT, conc)
        g.set_option(pkg.OPT_USE_GRAPH, graph)
        fc.setup_jet(g, case)
        for _ in range(2):
            g.flow_time_advance(dt)
        g.synchronize()
        out.append({n: g.download(n) for n in ("F", "DI", "T", "P")})
        g.close()
    for k, o in enumerate(out[1:], 1):
        for n in o:
            assert np.array_equal(o[n].view(np.uint64), out[0][n].view(np.uint64)), (k, n)


def test_three_steps_graph(pkg):
    case = fc.small_jet()
    g, o = pair(pkg, case)
    g.set_option(pkg.OPT_USE_GRAPH, 1)
    dt = case[5]["dt"]
    for _ in range(3):
        g.flow_time_advance(dt)
        o.flow_time_advance(dt)
    compare_all(g, o, tol=1e-11)
    g.close()


@pytest.mark.parametrize("mode", [1, -1])
def test_ns_step_with_chemistry(pkg, mode):
    import oracle
    from ocfd_amd import chemistry
    case = fc.small_jet()
    mech = case[5]["mech"]
    chemistry.ChemBatch(mech)            # mechanism into the HIP library
    oracle.ChemOracle(pkg, mech)         # and into the oracle
    g, o = pair(pkg, case)
    dt = case[5]["dt"]
    g.NS_step(dt, nsub=5, mode=mode)
    o.NS_step(dt, nsub=5, mode=mode)
    compare_all(g, o, tol=1e-11)
    g.close()


def test_ns_step_adaptive_chemistry(pkg):
    """NS_solver step with the adaptive point-implicit integrator in the
    loop against the oracle composed from its pieces: flow dt/2, the adaptive
    batch on the interior cells (f5 += dE), comput_T_P, flow dt/2."""
    import oracle
    from ocfd_amd import chemistry
    case = fc.small_jet()
    mech = case[5]["mech"]
    chemistry.ChemBatch(mech)
    co = oracle.ChemOracle(pkg, mech)
    g, o = pair(pkg, case)
    dt = case[5]["dt"] * 2.0             # flow halves at the case's CFL
    ctl = (chemistry.RTOL_Y, chemistry.Y_FLOOR, chemistry.RTOL_T, 4000)
    g.set_chem_method(pkg.CHEM_ADAPTIVE, *ctl, scheduler=2)
    g.NS_step(dt, nsub=1)
    o.flow_time_advance(0.5 * dt)
    di = o.download("DI")
    et = o.download("ET")
    f = o.download("F")
    ns = case[0].n_spec
    rho = np.stack([fc.interior(di[..., m]).reshape(-1, order="F") for m in range(ns)], axis=1)
    e0 = fc.interior(et).reshape(-1, order="F")
    r1, e1, nsub = co.adaptive_batch(rho, e0, np.full(e0.shape, dt), *ctl)
    print("substeps per cell: max", nsub.max(), "mean", nsub.mean())
    shp = fc.interior(et).shape
    for m in range(ns):
        fc.interior(di[..., m])[...] = r1[:, m].reshape(shp, order="F")
    f[..., 4] = f[..., 4] + (e1 - e0).reshape(shp, order="F")
    fc.interior(et)[...] = e1.reshape(shp, order="F")
    o.upload("DI", di)
    o.upload("ET", et)
    o.upload("F", f)
    o.comput_T_P()
    o.flow_time_advance(0.5 * dt)
    compare_all(g, o, tol=1e-10)
    g.close()


def test_rk3_step_with_ghost_cells(pkg):
    """Second-order ghost cells on the jet's x and y faces (species ghosts
    rescaled to d, real-gas Et and p) read by WENO5-with-ghost closures."""
    case = list(fc.small_jet())
    for a in range(4):
        case[0].ghost_cell[a] = 2
        case[0].scheme_boundary[a] = 1
    g, o = pair(pkg, tuple(case))
    dt = case[5]["dt"]
    g.flow_time_advance(dt)
    o.flow_time_advance(dt)
    compare_all(g, o)
    g.close()


def test_per_species_entries_equal_fused_step(pkg):
    """The reference's species loop through the per-species entries
    (du_scaler_invis, du_scaler_viscous, Scalar_time_adv_RK3 for m = 1..N_SPEC,
    OCFD_Time_Adv.f90:64-71) gives the fused step bit for bit; after the
    first species' pieces, Dus(1) and the partial Qp equal the oracle's."""
    import oracle
    case = fc.small_jet()
    dt = case[5]["dt"]
    ns = case[0].n_spec
    a = pkg.FlowSolver(case[0])
    fc.setup_jet(a, case)
    a.flow_time_advance(dt)
    fa = {n: a.download(n) for n in ("F", "DI", "T", "P")}
    a.close()
    o = oracle.Oracle(pkg, case[0])
    fc.setup_jet(o, case)
    o.comput_Amu_Amk_AMD()
    o.exchange_all()
    dus_o, _ = o.du_species()
    b = pkg.FlowSolver(case[0])
    fc.setup_jet(b, case)
    b.save_stage0()
    for krk in (1, 2, 3):
        b.stage_begin(krk)
        for m in range(1, ns + 1):
            b.du_scaler_invis(m)
            b.du_scaler_viscous(m)
            if krk == 1:
                dus = b.download("DUS")[..., m - 1]
                d, same = fc.compare(dus, dus_o[..., m - 1])
                assert d <= 1e-12, (m, d)
            b.scalar_time_adv_rk3(krk, m, dt)
        b.du_inviscous()
        b.du_viscous()
        b.OCFD_time_adv_RK3(krk, dt)
        b.comput_flow_variables()
        b.OCFD_bc()
    for n, v in fa.items():
        assert np.array_equal(v.view(np.uint64), b.download(n).view(np.uint64)), n
    with pytest.raises(pkg.OcfdError):
        b.du_scaler_invis(ns + 1)
    with pytest.raises(pkg.OcfdError):
        b.scalar_time_adv_rk3(4, 1, dt)
    b.close()
