"""Pins the chemistry oracle (CHEM_reactions.f90 / CHEM_thermdyn.f90
restatement) against conservation laws and independent numpy formulas."""
import math

import numpy as np
import pytest

from conftest import load_package

R0, ATM = 8.314, 1.013e5


@pytest.fixture(scope="module")
def chem():
    m = load_package()
    import oracle
    from ocfd_amd import mechanism
    M = mechanism.Mechanism()
    return oracle.ChemOracle(m, M), M


def mixture(M, T=1100.0, p=ATM, h2=0.3):
    X = np.zeros(M.n_spec)
    X[M.index("O2")] = 0.21 * (1 - h2)
    X[M.index("N2")] = 0.79 * (1 - h2)
    X[M.index("H2")] = h2
    Y = X * M.mi / np.sum(X * M.mi)
    rho = p / (T * np.sum(Y * R0 / M.mi))
    return rho * Y


def test_thermo_round_trip_and_continuity(chem):
    c, M = chem
    di = mixture(M)
    for T in (400.0, 999.0, 1001.0, 2500.0, 4999.0, 5200.0):
        E = c.E(di, T)
        assert abs(c.T(di, E) - T) < 1e-6
    # internal energy is continuous across Tct = 1000 K (Ect offsets)
    e_lo, e_hi = c.E(di, 1000.0 - 1e-9), c.E(di, 1000.0 + 1e-9)
    assert abs(e_hi - e_lo) < 1e-6 * abs(e_lo)


def nasa_gibbs_over_R(coef, T):
    a = coef
    H = a[0] * T + a[1] * T ** 2 / 2 + a[2] * T ** 3 / 3 + a[3] * T ** 4 / 4 + a[4] * T ** 5 / 5 + a[5]
    S = a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return H - T * S


@pytest.mark.parametrize("T", [800.0, 1500.0, 2400.0])
def test_equilibrium_constants_from_nasa_polynomials(chem, T):
    c, M = chem
    rng = (1 if T > M.tct[0] else 0)
    g = np.array([nasa_gibbs_over_R(M.coef[k][rng], T) for k in range(M.n_spec)])
    for k in range(M.n_spec):
        assert abs(c.gibbs(k, T) - g[k]) <= 1e-9 * max(1.0, abs(g[k]))
    ci = mixture(M, T) / M.mi * 1e-6
    wf, wr, kjx = c.rate_coef(ci, T)
    Rs = R0 / ATM * 1e6
    nu = (M.vr - M.vf)
    kc = np.exp(-(nu @ g) / T) * (Rs * T) ** (-nu.sum(axis=1).astype(float))
    assert np.allclose(kjx, kc, rtol=1e-11)


def test_rate_constants_closed_forms(chem):
    c, M = chem
    T = 1300.0
    ci = mixture(M, T) / M.mi * 1e-6
    wf, wr, kjx = c.rate_coef(ci, T)
    for j in range(M.n_reac):
        A, b, E = M.arrhenius[j, 0]
        k1 = A * T ** b * math.exp(-E / T)
        ctrb = float(np.dot(ci, M.eff[j])) if M.third_body[j] else 1.0
        if M.af_type[j] == 0:
            kf = k1 * (ctrb if M.third_body[j] else 1.0)
        elif M.af_type[j] == 2:
            A2, b2, E2 = M.arrhenius[j, 1]
            kf = (k1 + A2 * T ** b2 * math.exp(-E2 / T)) * (ctrb if M.third_body[j] else 1.0)
        else:  # Troe with the reference's single-precision 0.67 / 1.27
            A2, b2, E2 = M.arrhenius[j, 1]
            kinf = A2 * T ** b2 * math.exp(-E2 / T)
            Pr = k1 / kinf * ctrb
            Fc = M.fc_troe[j]
            ac = -0.4 - float(np.float32(0.67)) * math.log10(Fc)
            an = 0.75 - float(np.float32(1.27)) * math.log10(Fc)
            x = (math.log10(Pr) + ac) / (an - 0.14 * (math.log10(Pr) + ac))
            F = 10.0 ** (math.log10(Fc) / (1 + x * x))
            kf = kinf * Pr / (1 + Pr) * F
        assert abs(wf[j] - kf) <= 1e-12 * abs(kf), j
        expect_r = kf / kjx[j]
        assert abs(wr[j] - expect_r) <= 1e-12 * abs(expect_r)


def test_gauss_matches_lapack(chem):
    c, _ = chem
    rng = np.random.default_rng(0)
    for _ in range(20):
        A = rng.standard_normal((9, 9)) + 3 * np.eye(9)
        b = rng.standard_normal(9)
        assert np.allclose(c.gauss(A, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("mode", [1, 2, 3, -1])
def test_conservation_per_step(chem, mode):
    c, M = chem
    di = mixture(M, 1200.0)
    E = c.E(di, 1200.0)
    h = M.coef[:, 0, 5] * R0 / M.mi    # chemical energy per unit mass (update_Et)
    elem0 = (di / M.mi) @ M.elements
    etot0 = E + float(np.dot(di, h))
    d, e = di.copy(), E
    dt = 2e-7 if mode == -1 else 2e-9
    for _ in range(300):
        d, e = c.timeadv(mode, d, e, dt)
    assert abs(d.sum() - di.sum()) <= 1e-13 * di.sum()
    elem = (d / M.mi) @ M.elements
    assert np.allclose(elem, elem0, rtol=1e-12, atol=1e-14)
    assert abs(e + float(np.dot(d, h)) - etot0) <= 1e-10 * abs(etot0)


def test_point_implicit_step_is_the_linearised_solve(chem):
    c, M = chem
    di = mixture(M, 1400.0)
    E = c.E(di, 1400.0)
    d0, e0 = c.timeadv(-1, di, E, 3e-8)      # move off the fresh state
    T = c.T(d0, e0)
    ci = d0 / M.mi * 1e-6
    ri, A = c.rates(ci, T)
    dt = 1e-6
    A1 = np.eye(M.n_spec) - A * M.mi[:, None] / M.mi[None, :] * dt
    dd = np.linalg.solve(A1, ri * M.mi * 1e6 * dt)
    d1, _ = c.timeadv(-1, d0, e0, dt)
    assert np.allclose(d1, d0 + dd, rtol=1e-10, atol=1e-16 * d0.sum())


def test_jacobian_mass_action_part(chem):
    # for a state where third bodies are held fixed, the Jacobian of the
    # mass-action rates equals finite differences (the reference neglects the
    # third-body and temperature derivatives; they are excluded here)
    c, M = chem
    T = 1500.0
    ci = mixture(M, T) / M.mi * 1e-6
    ci = ci + 1e-9
    ri, A = c.rates(ci, T)
    wf, wr, _ = c.rate_coef(ci, T)
    # analytic: sum_k (vr-vf)_i * (vf_kj wf_k prod / c_j - vr_kj wr_k prod / c_j)
    nu = (M.vr - M.vf)
    fwd = wf * np.prod(ci[None, :] ** M.vf, axis=1)
    rev = wr * np.prod(ci[None, :] ** M.vr, axis=1)
    A_ref = nu.T @ ((M.vf * fwd[:, None] - M.vr * rev[:, None]) / (ci[None, :] + 1e-20))
    assert np.allclose(A, A_ref, rtol=1e-11, atol=1e-12 * np.max(np.abs(A_ref)))
    assert np.allclose(ri, nu.T @ (fwd - rev), rtol=1e-11, atol=1e-12 * np.max(np.abs(ri)))


def test_detailed_balance_at_equilibrium(chem):
    c, M = chem
    di = mixture(M, 1500.0, h2=0.25)
    E = c.E(di, 1500.0)
    d, e = di, E
    for _ in range(3000):
        d, e = c.timeadv(-1, d, e, 1e-6)
    T = c.T(d, e)
    ci = d / M.mi * 1e-6
    wf, wr, _ = c.rate_coef(ci, T)
    fwd = wf * np.prod(ci[None, :] ** M.vf, axis=1)
    rev = wr * np.prod(ci[None, :] ** M.vr, axis=1)
    big = fwd > 1e-12 * fwd.max()
    assert np.all(np.abs(fwd - rev)[big] <= 1e-4 * fwd[big])
    assert T > 2000.0


def test_adaptive_converges_to_refined_fixed_substeps(chem):
    c, M = chem
    di = mixture(M, 1100.0)
    E = c.E(di, 1100.0)
    # start just before ignition so the step crosses the ignition front
    d0, e0 = di, E
    for _ in range(780):
        d0, e0 = c.timeadv(-1, d0, e0, 1e-7)
    dt = 2e-5
    ref, eref = c.batch(-1, d0[None, :], np.array([e0]), np.array([dt]), 20000)
    ada, eada, nsub = c.adaptive_batch(d0[None, :], np.array([e0]), np.array([dt]), 0.05, 1e-6, 0.01, 4000)
    coarse, ec = c.batch(-1, d0[None, :], np.array([e0]), np.array([dt]), 1)
    Tr, Ta, Tc = c.T(ref[0], eref[0]), c.T(ada[0], eada[0]), c.T(coarse[0], ec[0])
    assert nsub[0] > 10
    assert abs(Ta - Tr) < 0.2 * abs(Tc - Tr) and abs(Ta - Tr) / Tr < 5e-3


# ---------------------------------------------------------------------------
# Explicit and point-implicit integrators against the closed-form stability
# polynomials of a linear decay, and the reverse-rate models ReReac = 1, 2.
# ---------------------------------------------------------------------------
def decay_mechanism(k):
    """Two species of equal molar mass, one irreversible reaction A -> B with
    a constant rate k (b = E = 0, no third body; ReReac = 1 with A_r = 0):
    d[A]/dt = -k [A] exactly, whatever the temperature.  The formation
    enthalpies differ so that update_Et moves E."""
    load_package()
    from ocfd_amd import mechanism
    coefA = [3.5, 0.0, 0.0, 0.0, 0.0, -1000.0, 4.0]
    coefB = [3.5, 0.0, 0.0, 0.0, 0.0, -4000.0, 4.0]
    data = dict(species=["A", "B"],
                mechanism=dict(n_spec=2, n_reac=1, rereac=1, names=["A", "B"], reactions=[dict(
                    label="A=>B", af_type=0, third_body=0, fc_troe=0.0, arrhenius=[(k, 0.0, 0.0)],
                    reverse=(0.0, 0.0, 0.0), vf=[1, 0], vr=[0, 1], eff=None)]),
                thermo=dict(n_tct=1, tct=[], tct_max=5000.0, species=[
                    dict(name="A", mi=0.028, spec_flag=1, coef=[coefA]),
                    dict(name="B", mi=0.028, spec_flag=1, coef=[coefB])]))
    return mechanism.Mechanism(data=data)


STABILITY = {1: lambda z: 1.0 - z,
             2: lambda z: 1.0 - z + z * z / 2.0,
             3: lambda z: 1.0 - z + z * z / 2.0 - z ** 3 / 6.0,
             -1: lambda z: 1.0 / (1.0 + z)}


@pytest.mark.parametrize("mode", [1, 2, 3, -1])
@pytest.mark.parametrize("z", [0.05, 0.4, 0.9])
def test_linear_decay_stability_polynomials(mode, z):
    """chemical_Euler / RK2 / RK3 / Euler_implicit (CHEM_reactions.f90:395-602)
    on d[A]/dt = -k[A]: one step of h = z/k multiplies [A] by the method's
    stability polynomial R(z) (1-z; 1-z+z^2/2; 1-z+z^2/2-z^3/6; 1/(1+z)).
    A wrong stage weight or a stage evaluated at the wrong state changes R at
    O(z^2).  The mass is conserved and E follows update_Et:
    E1 - E0 = -(dA R_A F_A + dB R_B F_B) with dB = -dA."""
    pkg = load_package()
    import oracle
    k = 2.0e5
    M = decay_mechanism(k)
    c = oracle.ChemOracle(pkg, M)
    di = np.array([0.7, 0.3])
    E0 = c.E(di, 1100.0)
    d1, E1 = c.timeadv(mode, di, E0, z / k)
    r = STABILITY[mode](z)
    assert abs(d1[0] - 0.7 * r) <= 1e-14 * 0.7, (d1[0], 0.7 * r)
    assert abs(d1.sum() - 1.0) <= 1e-15
    dA = d1[0] - 0.7
    R = R0 / 0.028
    expect = E0 - (dA * R * -1000.0 + (-dA) * R * -4000.0)
    assert abs(E1 - expect) <= 1e-12 * abs(E0)
    # substeps compose: n steps of z/n give R(z/n)^n
    dn, _ = c.batch(mode, di[None, :], np.array([E0]), np.array([z / k]), 8)
    assert abs(dn[0, 0] - 0.7 * STABILITY[mode](z / 8) ** 8) <= 1e-13 * 0.7


def _jet_with_reverse(rereac, rev):
    load_package()
    from ocfd_amd import mechanism
    M = mechanism.Mechanism()
    M.rereac = rereac
    M.reverse = np.asarray(rev, dtype=np.float64)
    return M


def _third_body(M, ci, j):
    return float(np.dot(ci, M.eff[j])) if M.third_body[j] else 1.0


def test_reverse_rates_arrhenius():
    """ReReac = 1: wr = A_r T^b_r exp(-E_r/T), times [M] for third-body
    reactions that are not fall-off (CHEM_reactions.f90:180-181, 190-195);
    the forward rates are those of ReReac = 0."""
    pkg = load_package()
    import oracle
    rng = np.random.default_rng(3)
    base = _jet_with_reverse(0, np.zeros((19, 5)))
    rev = np.zeros((19, 5))
    rev[:, 0] = 10.0 ** rng.uniform(8, 14, 19)
    rev[:, 1] = rng.uniform(-1.0, 1.0, 19)
    rev[:, 2] = rng.uniform(0.0, 3e4, 19)
    M = _jet_with_reverse(1, rev)
    T = 1700.0
    ci = mixture(M, T) / M.mi * 1e-6
    wf0, _, _ = oracle.ChemOracle(pkg, base).rate_coef(ci, T)
    wf, wr, _ = oracle.ChemOracle(pkg, M).rate_coef(ci, T)
    assert np.array_equal(wf, wf0)
    for j in range(M.n_reac):
        A, b, E = rev[j, :3]
        expect = A * T ** b * math.exp(-E / T)
        if M.third_body[j] and M.af_type[j] != 1:
            expect *= _third_body(M, ci, j)
        assert abs(wr[j] - expect) <= 1e-13 * expect, j


def test_reverse_rates_park_fit():
    """ReReac = 2: Kc = exp(A1 + A2 z + A3 z^2 + A4 z^3 + A5 z^4), z = 10000/T,
    floored at 1e-30, wr = wf / Kc (CHEM_reactions.f90:113-118, 182-186).
    Reactions 0..9 get coefficients whose Kc falls below the floor at this T."""
    pkg = load_package()
    import oracle
    rng = np.random.default_rng(4)
    rev = rng.uniform(-1.0, 1.0, (19, 5)) * np.array([5.0, 2.0, 0.5, 0.05, 0.005])
    rev[:10, 0] = -120.0
    rev[:10, 1:] *= 0.1
    M = _jet_with_reverse(2, rev)
    base = _jet_with_reverse(0, np.zeros((19, 5)))
    for T in (900.0, 2100.0):
        ci = mixture(M, T) / M.mi * 1e-6
        wf0, _, _ = oracle.ChemOracle(pkg, base).rate_coef(ci, T)
        wf, wr, kc = oracle.ChemOracle(pkg, M).rate_coef(ci, T)
        assert np.array_equal(wf, wf0)
        z = 1e4 / T
        for j in range(M.n_reac):
            a = rev[j]
            k = math.exp(a[0] + a[1] * z + a[2] * z ** 2 + a[3] * z ** 3 + a[4] * z ** 4)
            k = max(k, 1e-30)
            assert abs(kc[j] - k) <= 1e-13 * k, j
            assert abs(wr[j] - wf[j] / k) <= 1e-13 * abs(wf[j] / k), j
        assert np.all(kc[:10] == 1e-30)


def test_park_fit_of_gibbs_constants_reproduces_reref0():
    """Park coefficients fitted (least squares in ln Kc over 800..2800 K) to
    the Gibbs equilibrium constants of ReReac = 0 give nearly the same
    reverse rates: ties the units of the Park path (mol-cm, z = 10000/T) to
    the Gibbs path."""
    pkg = load_package()
    import oracle
    base = _jet_with_reverse(0, np.zeros((19, 5)))
    c0 = oracle.ChemOracle(pkg, base)
    Ts = np.linspace(800.0, 2800.0, 60)
    ci = mixture(base, 1500.0) / base.mi * 1e-6
    lk = np.array([np.log(c0.rate_coef(ci, T)[2]) for T in Ts])     # [T][reac]
    z = 1e4 / Ts
    V = np.stack([z ** p for p in range(5)], axis=1)
    rev = np.linalg.lstsq(V, lk, rcond=None)[0].T                    # [reac][5]
    c2 = oracle.ChemOracle(pkg, _jet_with_reverse(2, rev))
    for T in (1000.0, 1900.0):
        _, wr0, _ = c0.rate_coef(ci, T)
        _, wr2, _ = c2.rate_coef(ci, T)
        assert np.allclose(wr2, wr0, rtol=2e-3)
