"""GPU parity of the reacting planar H2 jet (WENO5, periodic span, BC_JET
with the subsonic-pressure inflow, 9 species / 19 reactions, real gas,
Chemkin/Wilke transport, mixture-averaged diffusion) and of the learned
surrogate deployed inside its NS_solver step.

Bars:
  * one RK3 step, and one NS_solver step with the reference's fixed-substep
    point-implicit chemistry (CHEM_TimeAdv = -1, 5 substeps): <= 1e-12
    relative (to each field's largest magnitude) per conserved variable,
    primitive and partial density against the oracle.  Differences come only
    from the device exp/log of the transport fits and rates (a few ulps per
    call); the fraction of bit-identical values is printed;
  * the surrogate inside the flow step (CHEM_SURROGATE): the step equals,
    bit for bit, the same step composed from its pieces (flow dt/2, the
    surrogate batch with fallback on the interior cells, f5 += dE,
    comput_T_P, flow dt/2); in that batch the cells sent to the fallback
    equal the plain adaptive FP64 integrator bit for bit, and the accepted
    cells conserve mass, H/O/N and E + sum_k rho_k R_k F_k to 1e-12.
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


def planar(**kw):
    load_package()
    from ocfd_amd import cases
    return cases.planar_jet(32, 24, 16, **kw)


def pair(pkg, case):
    import oracle
    g = pkg.FlowSolver(case[0])
    o = oracle.Oracle(pkg, case[0])
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


def test_planar_jet_case_shape():
    case = planar()
    cfg = case[0]
    assert (cfg.scheme_invis, cfg.n_spec, tuple(cfg.periodic), cfg.bc_para[0]) == (1, 9, (0, 0, 1), 1.0)


@pytest.mark.parametrize("iflag", [1, 0])
def test_planar_jet_rk3_step(pkg, iflag):
    case = planar(iflag=iflag)
    g, o = pair(pkg, case)
    dt = case[5]["dt"]
    g.flow_time_advance(dt)
    o.flow_time_advance(dt)
    compare_all(g, o)
    assert g.negative_et_count() == o.negative_et_count()
    g.close()


def test_planar_jet_species_right_hand_sides(pkg):
    """du_scaler_invis + du_scaler_viscous with the WENO5 reconstruction
    (interior and boundary closures) on a periodic span."""
    g, o = pair(pkg, planar())
    g.comput_Amu_Amk_AMD()
    o.comput_Amu_Amk_AMD()
    o.exchange_all()
    g.du_species()
    dus, qp = o.du_species()
    for name, a, b in (("DUS", g.download("DUS"), dus), ("QP", g.download("QP"), qp)):
        for m in range(a.shape[3]):
            d, same = fc.compare(a[..., m], b[..., m], name)
            print(name, m, f"{d:.1e}", f"{same:.3f}")
            assert d <= TOL, (name, m, d)
    g.close()


def test_planar_jet_ns_step_point_implicit(pkg):
    import oracle
    from ocfd_amd import chemistry
    case = planar()
    mech = case[5]["mech"]
    chemistry.ChemBatch(mech)
    oracle.ChemOracle(pkg, mech)
    g, o = pair(pkg, case)
    dt = case[5]["dt"]
    g.NS_step(dt, nsub=5, mode=-1)
    o.NS_step(dt, nsub=5, mode=-1)
    compare_all(g, o)
    g.close()


@pytest.fixture(scope="module")
def surrogate(pkg):
    """A briefly trained network (the checks under test do not depend on its
    accuracy) with its acceptance domain from the training cells."""
    import torch
    from ocfd_amd import chemistry, surrogate as sm, mechanism
    mech = mechanism.Mechanism()
    ch = chemistry.ChemBatch(mech)
    S = sm.Surrogate(ch, seed=0)
    data = S.dataset(1 << 17, seed=11)
    S.configure(data)
    x, y = S.training_arrays(data)
    S.train(x, y, epochs=3, batch=4096, lr=1e-3)
    torch.cuda.synchronize()
    return S, ch, mech


def _interior_cells(eng, ns):
    di, et = eng.download("DI"), eng.download("ET")
    rho = np.stack([fc.interior(di[..., m]).reshape(-1, order="F") for m in range(ns)], axis=0)
    return np.ascontiguousarray(rho), fc.interior(et).reshape(-1, order="F").copy()


def test_surrogate_in_flow_step(pkg, surrogate):
    import torch
    from ocfd_amd import chemistry
    S, ch, mech = surrogate
    case = planar()
    ns = case[0].n_spec
    dt = case[5]["dt"]
    ctl = (chemistry.RTOL_Y, chemistry.Y_FLOOR, chemistry.RTOL_T, 4000)
    # the trained domain reaches the cold jet stream; a floor of 900 K sends
    # the jet's 300 K cells to the fallback so both paths are exercised
    saved = S.cfg.copy()
    cfg = saved.copy()
    cfg[S.T_LO_INDEX] = 900.0
    S.set_config(cfg)
    try:
        _surrogate_in_flow_step(pkg, S, ch, mech, case, ns, dt, ctl)
    finally:
        S.set_config(saved)


def _surrogate_in_flow_step(pkg, S, ch, mech, case, ns, dt, ctl):
    import torch
    a = pkg.FlowSolver(case[0])
    fc.setup_jet(a, case)
    a.set_chem_method(pkg.CHEM_SURROGATE, *ctl, scheduler=2)
    a.NS_step(dt, nsub=1)
    a.synchronize()
    fused = {n: a.download(n) for n in ("F", "DI", "T", "P", "ET")}
    a.close()

    b = pkg.FlowSolver(case[0])
    fc.setup_jet(b, case)
    b.flow_time_advance(0.5 * dt)
    b.synchronize()
    rho0, e0 = _interior_cells(b, ns)
    dev = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda").contiguous()
    rho, E, dts = dev(rho0), dev(e0), dev(np.full(e0.shape, dt))
    flags = torch.zeros(e0.shape[0], dtype=torch.int32, device="cuda")
    S.step(rho, E, dts, fallback=True, controller=ctl[:3], max_sub=ctl[3], flags=flags)
    rp, Ep = dev(rho0), dev(e0)
    ch.adaptive(rp, Ep, dts, *ctl, scheduler=2)
    torch.cuda.synchronize()
    r1, e1, fl = rho.cpu().numpy(), E.cpu().numpy(), flags.cpu().numpy()
    fb = fl != 0
    print(f"cells {fl.size}: accepted {int(np.sum(~fb))}, fallback {int(np.sum(fb))}, reasons",
          {bit: int(np.sum((fl & bit) != 0)) for bit in (1, 2, 4, 8, 16)})
    assert fb.any() and (~fb).any(), "the case should exercise both the network and the fallback"
    # fallback cells: the plain adaptive integrator, bit for bit
    assert np.array_equal(r1[:, fb], rp.cpu().numpy()[:, fb])
    assert np.array_equal(e1[fb], Ep.cpu().numpy()[fb])
    # accepted cells: mass, elements and E + sum_k rho_k R_k F_k conserved
    acc = ~fb
    m0 = rho0[:, acc].sum(axis=0)
    assert np.max(np.abs(r1[:, acc].sum(axis=0) - m0) / m0) <= 1e-12
    el0 = (rho0[:, acc] / mech.mi[:, None]).T @ mech.elements
    el1 = (r1[:, acc] / mech.mi[:, None]).T @ mech.elements
    assert np.max(np.abs(el1 - el0) / (m0[:, None] / 0.001)) <= 1e-12
    hf = (mech.coef[:, 0, 5] * 8.314 / mech.mi)[:, None]
    t0 = e0[acc] + (rho0[:, acc] * hf).sum(axis=0)
    t1 = e1[acc] + (r1[:, acc] * hf).sum(axis=0)
    assert np.max(np.abs(t1 - t0) / (np.abs(e0[acc]) + np.abs(rho0[:, acc] * hf).sum(axis=0))) <= 1e-12

    # the fused NS step equals the composition bit for bit
    di, et, f = b.download("DI"), b.download("ET"), b.download("F")
    shp = fc.interior(et).shape
    for m in range(ns):
        fc.interior(di[..., m])[...] = r1[m].reshape(shp, order="F")
    f[..., 4] = f[..., 4] + (e1 - e0).reshape(shp, order="F")
    fc.interior(et)[...] = e1.reshape(shp, order="F")
    b.upload("DI", di)
    b.upload("ET", et)
    b.upload("F", f)
    b.comput_T_P()
    b.flow_time_advance(0.5 * dt)
    b.synchronize()
    for n, v in fused.items():
        w = b.download(n)
        assert np.array_equal(v.view(np.uint64), w.view(np.uint64)), n
    b.close()
