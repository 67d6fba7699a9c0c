"""GPU parity of the chemistry kernels (through the C-ABI) against the oracle.

The kernels keep the reference's operation order with -ffp-contract=off;
exp/log/log10/pow come from the device math library instead of glibc, so the
two sides differ by a few ulps per transcendental and the differences are
carried through the stiff update.  Tolerances, stated per test:
  * species: |rho_gpu - rho_oracle| <= RTOL_RHO * rho_cell (cell density),
  * energy:  |E_gpu - E_oracle| <= RTOL_E * |E|,
  * substep counts of the adaptive integrator: identical for >= 99.9% of
    cells (a count can move by one where an acceptance test or the end of dt
    is met within rounding), and the same number of unfinished cells.
"""
import numpy as np
import pytest

from conftest import load_package, gpu_available
import chemcase

pytestmark = pytest.mark.gpu

RTOL_RHO = 1e-10
RTOL_E = 1e-11


@pytest.fixture(scope="module")
def env():
    if not gpu_available():
        pytest.skip("no GPU")
    pkg = load_package()
    from ocfd_amd import chemistry
    b, M, orc = chemcase.jet_batch(1000)
    return pkg, chemistry.ChemBatch(M), b, M, orc


def to_dev(*arrs):
    import torch
    return [torch.tensor(a, dtype=torch.float64, device="cuda").contiguous() for a in arrs]


def report(tag, rg, eg, ro, eo):
    rho_cell = ro.sum(axis=0)
    dr = np.max(np.abs(rg - ro) / rho_cell[None, :])
    de = np.max(np.abs(eg - eo) / np.abs(eo))
    print(f"{tag}: max |drho|/rho_cell = {dr:.3e}, max |dE|/|E| = {de:.3e}, "
          f"bit-identical fraction = {np.mean(np.all(rg == ro, axis=0) & (eg == eo)):.4f}")
    return dr, de


@pytest.mark.parametrize("mode,nsub,dt_scale", [(-1, 1, 1.0), (-1, 5, 1.0), (1, 2, 1e-3), (2, 2, 1e-3), (3, 2, 1e-3)])
def test_fixed_substeps_match_oracle(env, mode, nsub, dt_scale):
    """chemical_timeAdv over nsub substeps (explicit modes at dt/1000, within
    their stability limit for these states)."""
    pkg, chem, b, M, orc = env
    dt = b["dt"] * dt_scale
    rho, E, d = to_dev(b["rho"], b["E"], dt)
    chem.chemical_timeAdv(rho, E, d, mode=mode, nsub=nsub)
    chem.synchronize()
    ro, eo = orc.batch(mode, b["rho"].T, b["E"], dt, nsub)
    dr, de = report(f"mode {mode} nsub {nsub}", rho.cpu().numpy(), E.cpu().numpy(), ro.T, eo)
    assert dr <= RTOL_RHO and de <= RTOL_E


@pytest.mark.parametrize("scheduler", [0, 1, 2])
def test_adaptive_matches_oracle(env, scheduler):
    pkg, chem, b, M, orc = env
    rho, E, d = to_dev(b["rho"], b["E"], b["dt"] * 20.0)
    ns = chem.adaptive(rho, E, d, rtol_y=0.05, y_floor=1e-6, rtol_t=0.01, max_sub=2000, scheduler=scheduler)
    chem.synchronize()
    ro, eo, no, uo = orc.adaptive_batch(b["rho"].T, b["E"], b["dt"] * 20.0, 0.05, 1e-6, 0.01, 2000,
                                        with_unfinished=True)
    assert chem.unfinished() == int(uo.sum())
    ng = ns.cpu().numpy()
    same = ng == no
    print(f"scheduler {scheduler}: substeps mean {no.mean():.1f} max {no.max()}, identical counts {same.mean():.4f}")
    assert same.mean() >= 0.999
    dr, de = report(f"adaptive sched {scheduler}", rho.cpu().numpy()[:, same], E.cpu().numpy()[same],
                    ro.T[:, same], eo[same])
    assert dr <= 1e3 * RTOL_RHO and de <= 1e3 * RTOL_E


def test_schedulers_bitwise_identical(env):
    """One cell per thread, the work queue and the work queue with the
    cooperative kernel run the same per-cell arithmetic: identical results.
    The long steps and a per-lane budget of 16 attempts make some cells move
    to the cooperative kernel."""
    pkg, chem, b, M, orc = env
    out = []
    for sched in (0, 1, 2):
        rho, E, d = to_dev(b["rho"], b["E"], b["dt"] * 50.0)
        chem.set_option(1, 16)
        try:
            ns = chem.adaptive(rho, E, d, max_sub=3000, scheduler=sched)
        finally:
            chem.set_option(1, 128)
        cnt = chem.counters()
        out.append((rho.cpu().numpy(), E.cpu().numpy(), ns.cpu().numpy()))
        if sched == 2:
            print("cells moved to the cooperative kernel:", cnt["cooperative"])
            assert cnt["cooperative"] > 0
    for o in out[1:]:
        for a, c in zip(out[0], o):
            assert np.array_equal(a, c)


def test_longest_first_order_is_bitwise_identical(env):
    """Schedulers 1 and 2 take the cells by decreasing dt (OPT_ORDER = 1,
    the default) or in index order (0): the same cells, the same per-cell
    arithmetic, so the same bits and substep counts, on a batch with dt over
    two decades (ties included) and on one with a single dt."""
    pkg, chem, b, M, orc = env
    from ocfd_amd import chemistry
    rng = np.random.default_rng(4)
    dt_var = b["dt"] * 10.0 ** rng.uniform(0.0, 2.0, b["dt"].shape[0])
    dt_var[::7] = dt_var[0]
    for dts in (dt_var, np.full_like(b["dt"], b["dt"].max())):
        for sched in (1, 2):
            out = []
            for order in (0, 1):
                chem.set_option(chemistry.OPT_ORDER, order)
                try:
                    rho, E, d = to_dev(b["rho"], b["E"], dts)
                    ns = chem.adaptive(rho, E, d, max_sub=3000, scheduler=sched)
                finally:
                    chem.set_option(chemistry.OPT_ORDER, 1)
                out.append((rho.cpu().numpy(), E.cpu().numpy(), ns.cpu().numpy(), chem.counters()["unfinished"]))
            for a, c in zip(out[0], out[1]):
                assert np.array_equal(a, c)
    with pytest.raises(pkg.OcfdError):
        chem.set_option(chemistry.OPT_ORDER, 2)


def test_unfinished_cells_are_counted(env):
    """A budget too small to cover dt leaves cells unfinished; both
    schedulers count them as the oracle does."""
    pkg, chem, b, M, orc = env
    dt = b["dt"] * 100.0
    ro, eo, no, uo = orc.adaptive_batch(b["rho"].T, b["E"], dt, 0.05, 1e-6, 0.01, 3, with_unfinished=True)
    assert 0 < uo.sum() < len(uo)
    for sched in (0, 1, 2):
        rho, E, d = to_dev(b["rho"], b["E"], dt)
        ns = chem.adaptive(rho, E, d, max_sub=3, scheduler=sched)
        assert chem.unfinished() == int(uo.sum())
        assert np.array_equal(ns.cpu().numpy(), no)


def test_ragged_and_edge_sizes(env):
    """n not a multiple of the workgroup, n = 1, n = 0, and cells with dt = 0
    (left untouched, zero substeps)."""
    import torch
    pkg, chem, b, M, orc = env
    for n in (1, 63, 129, 517):
        rho, E, d = to_dev(b["rho"][:, :n], b["E"][:n], b["dt"][:n])
        chem.chemical_timeAdv(rho, E, d, mode=-1, nsub=2)
        chem.synchronize()
        ro, eo = orc.batch(-1, b["rho"][:, :n].T, b["E"][:n], b["dt"][:n], 2)
        dr, de = report(f"n={n}", rho.cpu().numpy(), E.cpu().numpy(), ro.T, eo)
        assert dr <= RTOL_RHO and de <= RTOL_E
    empty = torch.zeros((M.n_spec, 0), dtype=torch.float64, device="cuda")
    e0 = torch.zeros(0, dtype=torch.float64, device="cuda")
    chem.chemical_timeAdv(empty, e0, e0.clone(), mode=-1)
    chem.adaptive(empty, e0, e0.clone())
    dt = b["dt"][:200].copy()
    dt[::3] = 0.0
    for sched in (0, 1, 2):
        rho, E, d = to_dev(b["rho"][:, :200], b["E"][:200], dt)
        ns = chem.adaptive(rho, E, d, scheduler=sched)
        chem.synchronize()
        r, e, n_ = rho.cpu().numpy(), E.cpu().numpy(), ns.cpu().numpy()
        assert np.all(n_[::3] == 0)
        assert np.array_equal(r[:, ::3], b["rho"][:, :200][:, ::3]) and np.array_equal(e[::3], b["E"][:200][::3])


def test_argument_errors(env):
    import torch
    pkg, chem, b, M, orc = env
    rho, E, d = to_dev(b["rho"][:, :8], b["E"][:8], b["dt"][:8])
    with pytest.raises(ValueError):
        chem.chemical_timeAdv(rho, E, d, mode=4)
    with pytest.raises(ValueError):
        chem.chemical_timeAdv(rho.t().contiguous(), E, d)
    with pytest.raises(TypeError):
        chem.chemical_timeAdv(rho.float(), E, d)
    with pytest.raises(pkg.OcfdError):
        chem.chemical_timeAdv(rho, E, d, mode=-1, nsub=0)
    with pytest.raises(pkg.OcfdError):
        chem.adaptive(rho, E, d, max_sub=0)


def test_conservation_at_scale(env):
    """Full-size property check (2^20 cells, generated and burnt on the GPU):
    per cell, total mass, H/O/N element moles and E + sum_k rho_k R_k F_k
    (update_Et's chemical energy) are invariant."""
    import torch
    pkg, chem, b, M, orc = env
    from ocfd_amd import cases
    n = 1 << 20
    s = cases.jet_chem_states(n, seed=11, mech=M)
    rho, E, tau = to_dev(s["rho"], s["E"], s["tau"])
    chem.adaptive(rho, E, tau, rtol_y=0.05, y_floor=1e-6, rtol_t=0.01, max_sub=100000)
    assert chem.unfinished() == 0
    r0, e0 = rho.clone(), E.clone()
    d = torch.tensor(s["dt"], dtype=torch.float64, device="cuda")
    nsub = chem.adaptive(rho, E, d, max_sub=1000)
    assert chem.unfinished() == 0
    # the longest-first order (default) against index order, at full size
    from ocfd_amd import chemistry
    rho_i, E_i = r0.clone(), e0.clone()
    chem.set_option(chemistry.OPT_ORDER, 0)
    try:
        nsub_i = chem.adaptive(rho_i, E_i, d, max_sub=1000)
    finally:
        chem.set_option(chemistry.OPT_ORDER, 1)
    assert torch.equal(rho, rho_i) and torch.equal(E, E_i) and torch.equal(nsub, nsub_i)
    torch.cuda.synchronize()
    mi = torch.tensor(M.mi, dtype=torch.float64, device="cuda")[:, None]
    el = torch.tensor(M.elements, dtype=torch.float64, device="cuda")
    hf = torch.tensor(M.coef[:, 0, 5] * 8.314 / M.mi, dtype=torch.float64, device="cuda")[:, None]
    m0, m1 = r0.sum(0), rho.sum(0)
    assert torch.max(torch.abs(m1 - m0) / m0).item() <= 1e-13
    x0, x1 = (r0 / mi).t() @ el, (rho / mi).t() @ el
    assert torch.max(torch.abs(x1 - x0) / (m0[:, None] / 0.001)).item() <= 1e-12
    et0, et1 = e0 + (r0 * hf).sum(0), E + (rho * hf).sum(0)
    scale = torch.abs(e0) + (torch.abs(r0 * hf)).sum(0)
    assert torch.max(torch.abs(et1 - et0) / scale).item() <= 1e-12
    assert torch.isfinite(rho).all() and int(nsub.min()) >= 1


def _reverse_data(M, orc, rereac):
    """Reverse-rate data that reproduce the Gibbs equilibrium of this
    mechanism: ReReac = 2, a Park fit of ln Kc in z = 10000/T; ReReac = 1, an
    Arrhenius fit of k_r = k_f / Kc (without third bodies) over 800..2800 K."""
    Ts = np.linspace(800.0, 2800.0, 60)
    X = np.zeros(M.n_spec)
    X[M.index("O2")], X[M.index("N2")], X[M.index("H2")] = 0.15, 0.55, 0.3
    ci = X * 101300.0 / (8.314 * 1500.0) * 1e-6
    kc, kr = [], []
    for T in Ts:
        wf, wr, kj = orc.rate_coef(ci, T)
        ctrb = np.where((M.third_body == 1) & (M.af_type != 1), M.eff @ ci, 1.0)
        kc.append(kj)
        kr.append(wr / ctrb)
    kc, kr = np.log(np.array(kc)), np.log(np.array(kr))
    rev = np.zeros((M.n_reac, 5))
    if rereac == 2:
        V = np.stack([(1e4 / Ts) ** p for p in range(5)], axis=1)
        rev[:] = np.linalg.lstsq(V, kc, rcond=None)[0].T
    else:
        V = np.stack([np.ones_like(Ts), np.log(Ts), -1.0 / Ts], axis=1)
        c = np.linalg.lstsq(V, kr, rcond=None)[0].T
        rev[:, 0], rev[:, 1], rev[:, 2] = np.exp(c[:, 0]), c[:, 1], c[:, 2]
    return rev


@pytest.mark.parametrize("rereac", [1, 2])
def test_reverse_rate_models_match_oracle(env, rereac):
    """ReReac = 1 (Arrhenius reverse rates) and 2 (Park fit of Kc, floored at
    1e-30), Reaction_rate_coef CHEM_reactions.f90:176-188: the point-implicit
    integrator (5 substeps) and the adaptive one (scheduler 2, the lane queue
    plus the cooperative kernel) against the oracle with the same data."""
    import copy
    import oracle
    from ocfd_amd import chemistry
    pkg, chem, b, M, orc = env
    M2 = copy.deepcopy(M)
    M2.rereac = rereac
    M2.reverse = _reverse_data(M, orc, rereac)
    try:
        ch2 = chemistry.ChemBatch(M2)
        o2 = oracle.ChemOracle(pkg, M2)
        rho, E, d = to_dev(b["rho"], b["E"], b["dt"])
        ch2.chemical_timeAdv(rho, E, d, mode=-1, nsub=5)
        ch2.synchronize()
        ro, eo = o2.batch(-1, b["rho"].T, b["E"], b["dt"], 5)
        dr, de = report(f"ReReac {rereac} point-implicit", rho.cpu().numpy(), E.cpu().numpy(), ro.T, eo)
        assert dr <= RTOL_RHO and de <= RTOL_E
        rho, E, d = to_dev(b["rho"], b["E"], b["dt"] * 20.0)
        ns = ch2.adaptive(rho, E, d, rtol_y=0.05, y_floor=1e-6, rtol_t=0.01, max_sub=2000, scheduler=2)
        ch2.synchronize()
        ro, eo, no = o2.adaptive_batch(b["rho"].T, b["E"], b["dt"] * 20.0, 0.05, 1e-6, 0.01, 2000)
        same = ns.cpu().numpy() == no
        assert same.mean() >= 0.999
        dr, de = report(f"ReReac {rereac} adaptive", rho.cpu().numpy()[:, same], E.cpu().numpy()[same],
                        ro.T[:, same], eo[same])
        assert dr <= 1e3 * RTOL_RHO and de <= 1e3 * RTOL_E
    finally:
        chemistry.ChemBatch(M)          # the module's other tests use ReReac = 0
        oracle.ChemOracle(pkg, M)


def test_reverse_rate_data_required(env):
    """ReReac = 1 or 2 without reverse data: the batch calls refuse."""
    import copy
    from ocfd_amd import chemistry, OcfdError
    pkg, chem, b, M, orc = env
    lib = chem.lib
    k = chem._keep
    try:
        rc = lib.ocfd_chem_set_mechanism(M.n_spec, M.n_reac, 2, *[chemistry._vp(k[n]) for n in
                                         ("af", "tb", "fc", "arr", "vf", "vr", "eff")], M.n_tct,
                                         chemistry._vp(k["tct"]), float(M.tct_max), chemistry._vp(k["mi"]),
                                         chemistry._vp(k["coef"]))
        assert rc == 0
        rho, E, d = to_dev(b["rho"][:, :8], b["E"][:8], b["dt"][:8])
        with pytest.raises(OcfdError, match="reverse"):
            chem.chemical_timeAdv(rho, E, d, mode=-1, nsub=1)
    finally:
        chemistry.ChemBatch(M)
