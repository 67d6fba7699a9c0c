"""Host-side pieces of the learned chemistry surrogate, on the CPU: the
element-balance projection and the calibrated increment bound."""
import numpy as np

from conftest import load_package


def sur():
    load_package()
    from ocfd_amd import surrogate, mechanism
    return surrogate, mechanism.Mechanism()


def test_projection_conserves_elements_and_is_a_projector():
    s, M = sur()
    cel = s.element_matrix(M)
    rng = np.random.default_rng(0)
    sigma = 10 ** rng.uniform(-6, -1, cel.shape[1])
    P = s.projection(cel, sigma)
    dY = rng.standard_normal((50, cel.shape[1])) * sigma
    # C P dY = 0 for every increment: H, O and N are conserved
    assert np.max(np.abs(cel @ P @ dY.T)) <= 1e-12 * np.max(np.abs(cel @ dY.T))
    # idempotent, and increments that already conserve elements pass unchanged
    assert np.allclose(P @ P, P, atol=1e-12)
    good = P @ dY.T
    assert np.allclose(P @ good, good, rtol=0, atol=1e-15)
    # mass is a sum of element masses, so it is conserved too
    assert np.max(np.abs(good.sum(axis=0))) <= 1e-12 * np.max(np.abs(dY))


def brute_bound(m, err, tol):
    """Largest candidate bound (each m, and 0) that keeps every cell at or
    below it within tol."""
    best = 0.0
    for t in sorted(set(m.tolist())):
        if np.all(err[m <= t] <= tol):
            best = t
        else:
            break
    return best


def test_trust_bound_against_brute_force():
    s, _ = sur()
    rng = np.random.default_rng(1)
    for trial in range(200):
        n = int(rng.integers(1, 40))
        m = np.round(rng.uniform(0, 1, n), 2)          # ties on purpose
        err = rng.uniform(0, 1, n) * m * rng.uniform(0.5, 2.0)
        tol = rng.uniform(0.05, 0.5)
        b = s.trust_bound(m, err, tol)
        bf = brute_bound(m, err, tol)
        # every cell the bound keeps is within tol
        assert np.all(err[m <= b] <= tol), trial
        # and it keeps exactly the cells the largest admissible threshold keeps
        bad = err > tol
        if bad.any():
            assert np.array_equal(m <= b, m <= bf), trial
            assert b < m[bad].min()
        else:
            assert np.all(m <= b)


def test_trust_bound_cases():
    s, _ = sur()
    m = np.array([0.1, 0.2, 0.3, 0.4])
    # the cell at 0.3 errs: the cell at 0.4 is not trusted either
    assert 0.2 < s.trust_bound(m, np.array([1e-4, 3e-4, 2e-3, 1e-4]), 5e-4) < 0.3
    # nothing errs: just above the largest increment seen
    b = s.trust_bound(m, np.zeros(4), 5e-4)
    assert 0.4 < b < 0.4 * (1 + 1e-5)
    # the smallest increment errs already, or no cells: trust nothing
    assert s.trust_bound(m, np.array([1.0, 0, 0, 0]), 5e-4) == 0.0
    assert s.trust_bound(np.array([]), np.array([]), 5e-4) == 0.0
    # NaN errors count as errors
    assert 0.1 < s.trust_bound(m, np.array([0, np.nan, 0, 0]), 5e-4) < 0.2


def test_config_layout():
    s, _ = sur()
    # in_shift, in_scale (12 each), out_scale (9), domain (6), the acceptance
    # bounds y_neg_tol, elem_tol, T_out_lo, T_out_hi, dy_max, P (81), Cel (27)
    assert s.CFG_LEN == 12 + 12 + 9 + 6 + 5 + 81 + 27
    assert s.Surrogate.DY_MAX_INDEX == 12 + 12 + 9 + 6 + 4
