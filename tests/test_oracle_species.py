"""Pins the multi-species parts of the flow oracle (N_SPEC > 1): real-gas
thermodynamics, Chemkin/Wilke transport, the species convection/diffusion
and update, Rescaling_di, the jet boundary condition and the chemistry
coupling.  The reference ships no golden vectors for these, so each function
is checked against physical data, closed forms and invariants chosen so that
a plausible slip (wrong unit factor, swapped index, sign error, dropped
term) fails at least one of them.  CPU only.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import load_package
import flowcase as fc

LAP = 4
R0 = 8.314


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def mech():
    load_package()
    from ocfd_amd.mechanism import Mechanism
    return Mechanism()


def uniform_oracle(m, mech, X, T, p=101300.0, n=8, real_gas=1, flags=1):
    """An oracle context with N_SPEC = 9 and a uniform state of mole
    fractions X at T, p (interior n^3)."""
    import oracle
    cfg = m.make_cfg(n, n, n, periodic=(1, 1, 1), n_spec=mech.n_spec, if_real_gas=real_gas, flag_vis=flags,
                     flag_cond=flags, flag_diff=flags)
    o = oracle.Oracle(m, cfg)
    o.set_species(mech.mi, n_tct=mech.n_tct, tct=mech.tct, tct_max=mech.tct_max, coef=mech.coef)
    o.set_transport(mech.viscosity, mech.conductivity, mech.diffusion)
    Y = mech.mass_fractions(np.asarray(X, dtype=np.float64))
    d = p / (T * np.sum(Y * R0 / mech.mi))
    h = (n + 2 * LAP,) * 3
    di = np.ones(h + (mech.n_spec,), order="F")
    for k in range(mech.n_spec):
        di[..., k] = Y[k] * d
    o.upload("DI", di)
    o.upload("D", np.full(h, d, order="F"))
    o.upload("T", np.full(h, T, order="F"))
    o.upload("P", np.full(h, p, order="F"))
    return o, d, Y


def mole(mech, **xs):
    X = np.zeros(mech.n_spec)
    for k, v in xs.items():
        X[mech.index(k)] = v
    return X


# --------------------------------------------------------------- thermo
def test_nasa_cp_of_nitrogen_and_water(mech):
    m = load_package()
    o, d, Y = uniform_oracle(m, mech, mole(mech, N2=1.0), 300.0)
    di = np.zeros(mech.n_spec)
    di[mech.index("N2")] = 1.0
    # Cp(N2, 300 K) = 29.12 J/mol/K (JANAF) -> 1040 J/kg/K with M = 0.028
    cp = o.lib.orc_comput_Cp(1.0, P(di), 300.0)
    assert abs(cp - 29.12 / 0.028) / (29.12 / 0.028) < 5e-3
    di[:] = 0.0
    di[mech.index("H2O")] = 1.0
    # Cp(H2O, 1500 K) = 47.06 J/mol/K (JANAF), range 2 of the fit
    cp = o.lib.orc_comput_Cp(1.0, P(di), 1500.0)
    assert abs(cp - 47.06 / mech.mi[mech.index("H2O")]) / (47.06 / mech.mi[mech.index("H2O")]) < 1e-2
    o.close()


def test_energy_continuous_cp_consistent_and_invertible(mech):
    m = load_package()
    o, d, Y = uniform_oracle(m, mech, mole(mech, H2=0.3, O2=0.15, H2O=0.1, OH=0.01, N2=0.44), 1200.0)
    rng = np.random.default_rng(3)
    for trial in range(20):
        di = rng.uniform(0.0, 0.3, mech.n_spec)
        dd = di.sum()
        Rm = float(np.sum(di * R0 / mech.mi))
        tct = mech.tct[0]
        # E continuous across the range boundary (Ect offsets)
        e_lo = o.lib.orc_comput_E(P(di), tct * (1 - 1e-12))
        e_hi = o.lib.orc_comput_E(P(di), tct * (1 + 1e-12))
        assert abs(e_hi - e_lo) <= 1e-9 * abs(e_lo)
        for T in (400.0, 999.0, 1500.0, 3000.0, 5500.0):   # 5500 K: constant Cp beyond tct_max
            h = 1e-3 * T
            dE = (o.lib.orc_comput_E(P(di), T + h) - o.lib.orc_comput_E(P(di), T - h)) / (2 * h)
            cp = o.lib.orc_comput_Cp(dd, P(di), T)
            assert abs(cp - (dE + Rm) / dd) <= 1e-6 * cp, (trial, T)
            E = o.lib.orc_comput_E(P(di), T)
            assert abs(o.lib.orc_comput_T(P(di), E) - T) < 1e-3
    o.close()


# --------------------------------------------------------------- transport
def fit(a, T):
    lT = math.log(T)
    return math.exp(a[0] + a[1] * lT + a[2] * lT ** 2 + a[3] * lT ** 3)


def test_pure_species_transport_physical_values(mech):
    """Pure-gas limits of the Wilke / conductivity mixing rules against
    tabulated data at 300 K (units: the fits are g/cm/s and erg/cm/s/K)."""
    m = load_package()
    data = {"N2": (1.78e-5, 0.0259), "O2": (2.07e-5, 0.0266), "H2": (8.9e-6, 0.182)}
    for sp, (mu_ref, k_ref) in data.items():
        o, d, Y = uniform_oracle(m, mech, mole(mech, **{sp: 1.0}), 300.0)
        o.comput_Amu_Amk_AMD()
        mu, k = o.download("AMU"), o.download("AMK")
        assert np.all(mu == mu.flat[0]) and np.all(k == k.flat[0])
        assert abs(mu.flat[0] - mu_ref) / mu_ref < 0.03, (sp, mu.flat[0])
        assert abs(k.flat[0] - k_ref) / k_ref < 0.06, (sp, k.flat[0])
        j = mech.index(sp)
        assert abs(mu.flat[0] - 0.1 * fit(mech.viscosity[j], 300.0)) <= 1e-14 * mu.flat[0]
        o.close()


def test_binary_mixture_rules(mech):
    """H2/N2 binary: Wilke's two-species form, the 1/2 (sum X k + 1/sum X/k)
    conductivity rule and D_k = (1 - X_k) / (X_j / D_kj) atm/p; the binary
    diffusion coefficient of H2-N2 at 300 K, 1 atm is 0.78 cm^2/s."""
    m = load_package()
    T, p = 300.0, 2.0 * 101300.0
    x1 = 0.37
    o, d, Y = uniform_oracle(m, mech, mole(mech, H2=x1, N2=1 - x1), T, p=p)
    o.comput_Amu_Amk_AMD()
    i, j = mech.index("H2"), mech.index("N2")
    M1, M2 = mech.mi[i], mech.mi[j]
    mu1, mu2 = 0.1 * fit(mech.viscosity[i], T), 0.1 * fit(mech.viscosity[j], T)
    phi12 = (1 + math.sqrt(mu1 / mu2) * (M2 / M1) ** 0.25) ** 2 / math.sqrt(8 * (1 + M1 / M2))
    phi21 = (1 + math.sqrt(mu2 / mu1) * (M1 / M2) ** 0.25) ** 2 / math.sqrt(8 * (1 + M2 / M1))
    x2 = 1 - x1
    mu_w = x1 * mu1 / (x1 + x2 * phi12) + x2 * mu2 / (x2 + x1 * phi21)
    assert abs(o.download("AMU").flat[0] - mu_w) <= 1e-12 * mu_w
    k1, k2 = 1e-5 * fit(mech.conductivity[i], T), 1e-5 * fit(mech.conductivity[j], T)
    k_mix = 0.5 * (x1 * k1 + x2 * k2 + 1.0 / (x1 / k1 + x2 / k2))
    assert abs(o.download("AMK").flat[0] - k_mix) <= 1e-12 * k_mix
    amd = o.download("AMD")
    D12 = 1e-4 * fit(mech.diffusion[i][j], T)
    assert abs(D12 - 0.78e-4) / 0.78e-4 < 0.1
    for k, xk in ((i, x1), (j, x2)):
        expect = (1 - xk) / ((1 - xk) / D12) * 101300.0 / p
        assert abs(amd[..., k].flat[0] - expect) <= 1e-12 * expect
    o.close()


# --------------------------------------------------------------- species transport
def two_species_isotropic(a_field=None, n=16):
    """The isotropic case as a mixture of two identical species (air) with
    di_1 = a d, di_2 = (1 - a) d."""
    m = load_package()
    cfg, mesh, span, st, bc, info = fc.small_isotropic(n)
    cfg.n_spec = 2
    d = st["D"]
    a = 0.3 if a_field is None else a_field
    di = np.ones(d.shape + (2,), order="F")
    di[..., 0] = a * d
    di[..., 1] = d - di[..., 0]
    st = dict(st)
    st["DI"] = di
    return m, (cfg, mesh, span, st, bc, info)


def run_oracle(m, case, steps, ns_mi=None):
    import oracle
    o = oracle.Oracle(m, case[0])
    if case[0].n_spec > 1:
        o.set_species(np.full(case[0].n_spec, 0.02897))
    fc.setup(o, case)
    for _ in range(steps):
        o.flow_time_advance(case[5]["dt"])
    return o


def test_identical_species_reduce_to_single_species_flow():
    """Two identical species carry the same flow as N_SPEC = 1: d, u, v, w,
    T, p agree to round-off after two steps, whatever the composition."""
    m, case2 = two_species_isotropic()
    n = 16
    x = np.arange(n) * (2 * math.pi / n)
    X = np.meshgrid(x, x, x, indexing="ij")
    a = np.full((n + 2 * LAP,) * 3, 0.5)
    fc.interior(a)[...] = 0.5 + 0.4 * np.sin(X[0]) * np.cos(2 * X[1]) * np.sin(X[2] + 0.3)
    m, case2 = two_species_isotropic(a)
    o2 = run_oracle(m, case2, 2)
    res2 = {k: o2.download(k) for k in ("F", "D", "U", "T", "P")}
    di = o2.download("DI")
    o2.close()
    case1 = fc.small_isotropic(n)
    o1 = run_oracle(m, case1, 2)
    for k, v in res2.items():
        ref = o1.download(k)
        if k != "F":
            v, ref = fc.interior(v), fc.interior(ref)
        assert np.max(np.abs(v - ref)) <= 1e-12 * np.max(np.abs(ref)), k
    # the partial densities still sum to d (Rescaling_di)
    s = fc.interior(di[..., 0] + di[..., 1])
    assert np.max(np.abs(s - fc.interior(res2["D"]))) <= 1e-14 * np.max(s)
    o1.close()


def periodic_mixture(n, u, v, w, a, sc=1.0):
    """Two identical species (air) on the periodic 2 pi box: uniform d and
    T, the given velocity and first-species fraction fields (interior)."""
    m = load_package()
    from ocfd_amd import cases
    L = 2 * math.pi
    x = np.arange(n) * (L / n)
    d = np.full((n, n, n), 1.2)
    T = np.full((n, n, n), 300.0)
    st = cases.state_from_primitives(d, u, v, w, T)
    di = np.ones(st["D"].shape + (2,), order="F")
    fc.interior(di[..., 0])[...] = a * d
    fc.interior(di[..., 1])[...] = d - a * d
    st["DI"] = di
    cfg = m.make_cfg(n, n, n, periodic=(1, 1, 1), n_spec=2, sc=sc)
    mesh = cases.mesh_1d(x, x, x)
    span = np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)
    c0 = math.sqrt(1.4 * cases.R_AIR * 300.0)
    dt = 0.3 / float(np.max((np.abs(u) + c0) + (np.abs(v) + c0) + (np.abs(w) + c0)) / (L / n))
    return m, (cfg, mesh, span, st, None, dict(dt=dt))


def oracle_for(m, case):
    import oracle
    o = oracle.Oracle(m, case[0])
    o.set_species(np.full(case[0].n_spec, 0.02897))
    fc.setup(o, case)
    return o


def test_supersonic_species_flux_is_fraction_of_mass_flux():
    """Supersonic in every direction, Steger-Warming sends the whole mass
    flux upwind, as the scalar split does with the species flux: with a
    uniform fraction a, the species right-hand side is a times the density
    right-hand side, and the fraction stays a through RK3 steps.  A wrong
    split, direction, metric factor, sign or update fails this."""
    n = 16
    L = 2 * math.pi
    x = np.arange(n) * (L / n)
    X = np.meshgrid(x, x, x, indexing="ij")
    c0 = math.sqrt(1.4 * 287.0 * 300.0)
    u = 3.0 * c0 + 20.0 * np.sin(X[0]) * np.cos(X[2])
    v = 2.5 * c0 + 20.0 * np.sin(X[1] + 1.0) * np.cos(X[0])
    w = 2.0 * c0 + 20.0 * np.sin(X[2] + 2.0) * np.cos(X[1])
    m, case = periodic_mixture(n, u, v, w, 0.3)
    o = oracle_for(m, case)
    o.comput_Amu_Amk_AMD()
    o.exchange_all()
    dus, qp = o.du_species()
    o.du_inviscous()
    du = o.download("DU")
    scale = np.max(np.abs(du[..., 0]))
    assert scale > 0
    assert np.max(np.abs(dus[..., 0] - 0.3 * du[..., 0])) <= 1e-12 * scale
    assert np.max(np.abs(dus[..., 0] + dus[..., 1] - du[..., 0])) <= 1e-12 * scale
    o.close()
    o = oracle_for(m, case)
    for _ in range(2):
        o.flow_time_advance(case[5]["dt"])
    di, d = fc.interior(o.download("DI")), fc.interior(o.download("D"))
    assert np.max(np.abs(d - 1.2)) > 1e-6          # the density did change
    assert np.max(np.abs(di[..., 0] / d - 0.3)) < 1e-13
    o.close()


def test_species_diffusion_closed_form():
    """At rest with c = 0.3 + 0.1 sin(2x), the species right-hand side is
    rho D d2c/dx2 = -0.4 rho D sin(2x) with D = mu / (Sc rho) for two
    identical species; the diffusive enthalpy flux sum_k rho D_k h_k grad c_k
    cancels.  CD6 error falls by about 2^6 from n = 16 to 32."""
    errs = []
    for n in (16, 32):
        L = 2 * math.pi
        x = np.arange(n) * (L / n)
        X = np.meshgrid(x, x, x, indexing="ij")
        z = np.zeros((n, n, n))
        a = 0.3 + 0.1 * np.sin(2 * X[0])
        m, case = periodic_mixture(n, z, z, z, a, sc=0.7)
        o = oracle_for(m, case)
        o.comput_Amu_Amk_AMD()
        o.exchange_all()
        dus, qp = o.du_species()
        mu = o.download("AMU")
        amd = o.download("AMD")
        assert np.allclose(amd[..., 0], mu / (0.7 * 1.2), rtol=1e-13, atol=0)
        exact = -0.1 * 4.0 * mu / 0.7 * np.sin(2 * X[0])
        errs.append(np.max(np.abs(dus[..., 0] - exact)) / np.max(np.abs(exact)))
        assert np.max(np.abs(dus[..., 0] + dus[..., 1])) <= 1e-12 * np.max(np.abs(exact))
        assert np.max(np.abs(qp)) <= 1e-10 * np.max(np.abs(exact)) * 3.5 * 287.0 * 300.0
        o.close()
    assert errs[0] < 5e-3 and errs[0] / errs[1] > 40, errs


# --------------------------------------------------------------- jet BC + chemistry
def jet_oracle(iflag=1, n=(24, 16, 16)):
    m = load_package()
    import oracle
    case = fc.small_jet(*n, iflag=iflag)
    o = oracle.Oracle(m, case[0])
    fc.setup_jet(o, case)
    return m, o, case


def test_jet_inflow_and_outflow():
    m, o, case = jet_oracle(iflag=1)
    o.flow_time_advance(case[5]["dt"])
    p, T, u, d = (o.download(k) for k in ("P", "T", "U", "D"))
    di = o.download("DI")
    inl = case[4]["inlet2d"]
    mech = case[5]["mech"]
    ny, nz = inl.shape[:2]
    ai = inl[..., 5:] / inl[..., 5:].sum(axis=2, keepdims=True)
    Rm = (ai * (R0 / mech.mi)).sum(axis=2)
    p_in = (inl[..., 5:] * (R0 / mech.mi)).sum(axis=2) * inl[..., 4]
    E_in = mech.energy(inl[..., 5:].reshape(-1, 9).T, inl[..., 4].ravel()).reshape(ny, nz)
    c_in = np.sqrt((p_in / E_in + 1.0) * p_in / inl[..., 0])
    sub = inl[..., 1] < c_in
    assert sub.any() and (~sub).any()
    i1, i2, inx, inx1 = LAP, LAP + 1, LAP + case[0].nx - 1, LAP + case[0].nx - 2
    J = slice(LAP, LAP + ny)
    K = slice(LAP, LAP + nz)
    # subsonic: pressure from i = 2, density from p / (R_mix T); supersonic: data
    assert np.allclose(p[i1, J, K][sub], p[i2, J, K][sub], rtol=1e-14, atol=0)
    assert np.allclose(d[i1, J, K][sub], p[i2, J, K][sub] / (Rm[sub] * inl[..., 4][sub]), rtol=1e-14, atol=0)
    assert np.allclose(d[i1, J, K][~sub], inl[..., 0][~sub], rtol=1e-14, atol=0)
    assert np.array_equal(u[i1, J, K], inl[..., 1]) and np.array_equal(T[i1, J, K], inl[..., 4])
    # outflow: zeroth-order extrapolation
    for a in (d, u, T):
        assert np.array_equal(a[inx, J, K], a[inx1, J, K])
    # p and di pass through the face re-synthesis (di rescaled to sum to d)
    assert np.allclose(p[inx, J, K], p[inx1, J, K], rtol=1e-14, atol=0)
    assert np.allclose(di[inx, J, K], di[inx1, J, K], rtol=1e-14, atol=1e-14 * np.max(di))
    o.close()


def test_chemistry_coupling_conserves(mech):
    """Chemical_time_advance changes only the composition and the internal
    energy: d and the kinetic part f5 - Et are unchanged, elements are
    conserved and E + sum rho_k R_k F_k (sensible + formation) is conserved."""
    import oracle
    m, o, case = jet_oracle()
    oracle.ChemOracle(m, mech)
    f0, et0, di0 = o.download("F"), fc.interior(o.download("ET")), fc.interior(o.download("DI"))
    o.Chemical_time_advance(2e-7, 1)
    f1, et1, di1 = o.download("F"), fc.interior(o.download("ET")), fc.interior(o.download("DI"))
    assert np.max(np.abs(di1 - di0)) > 0
    assert np.allclose(di1.sum(axis=3), di0.sum(axis=3), rtol=1e-13, atol=0)
    assert np.allclose(f1[..., 4] - et1, f0[..., 4] - et0, rtol=1e-12, atol=1e-9 * np.max(np.abs(et0)))
    el0 = (di0 / mech.mi) @ mech.elements
    el1 = (di1 / mech.mi) @ mech.elements
    assert np.max(np.abs(el1 - el0)) <= 1e-12 * np.max(np.abs(el0))
    hf = mech.coef[:, 0, 5] * R0 / mech.mi
    t0, t1 = et0 + (di0 * hf).sum(axis=3), et1 + (di1 * hf).sum(axis=3)
    assert np.max(np.abs(t1 - t0)) <= 1e-12 * np.max(np.abs(et0) + np.abs(di0 * hf).sum(axis=3))
    o.close()
