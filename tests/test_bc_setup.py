"""The boundary-layer set-up routines on the host (cases.init_bc_nonref,
init_bc_wall_perturbation, get_ht_multifrequancy; OCFD_BC_Boundarylayers.f90
:406-443, :446-562, :594-621) against values their formulas fix: the
Dirichlet/non-reflecting split by the top surface's slope, the strip profile
fx whose peak is 1 at seta = 2 pi / 3, the spanwise factors for mz = 0,
mz < 0 and mz > 0, the random strip's range, and the temporal factor."""
import math

import numpy as np
import pytest

from conftest import load_package


@pytest.fixture(scope="module")
def cases():
    load_package()
    from ocfd_amd import cases
    return cases


def test_nonref_types(cases):
    nx, nz = 12, 4
    X = np.linspace(0.0, 1.1, nx)[:, None] * np.ones((1, nz))
    flat = np.full((nx, nz), 2.0)
    assert np.all(cases.init_bc_nonref(1, X, flat) == 1)
    assert np.all(cases.init_bc_nonref(2, X, flat) == 0)
    # auto: flat top -> non-reflecting; rising top -> Dirichlet; a top that
    # rises then falls splits at the turning point (backward slope)
    assert np.all(cases.init_bc_nonref(0, X, flat) == 1)
    assert np.all(cases.init_bc_nonref(0, X, 2.0 + 0.3 * X) == 0)
    bump = 2.0 + 0.3 * np.minimum(X, 0.5)
    nr = cases.init_bc_nonref(0, X, bump)
    x = X[:, 0]
    rising = np.r_[True, x[1:] <= 0.5 + 1e-12] & np.r_[True, x[:-1] < 0.5]
    assert np.array_equal(nr[:, 0] == 0, rising)
    # angle of attack: the reference compares against tan(AoA) + 1e-6
    slope = 0.3
    assert np.all(cases.init_bc_nonref(0, X, 2.0 + slope * X, aoa=math.atan(slope) + 1e-3) == 1)
    assert np.all(cases.init_bc_nonref(0, X, 2.0 + slope * X, aoa=math.atan(slope) - 1e-3) == 0)
    with pytest.raises(ValueError):
        cases.init_bc_nonref(3, X, flat)


def test_wall_perturbation_strip_and_modes(cases):
    nx, nz, ZL = 31, 16, 0.02
    xb, xe = 0.1, 0.4
    x = np.linspace(0.0, 0.5, nx)
    z = np.arange(nz) * (ZL / nz)
    X, Z = np.meshgrid(x, z, indexing="ij")
    rng = np.random.default_rng(1)
    assert np.all(cases.init_bc_wall_perturbation(0, xb, xe, 5, ZL, X, Z, rng) == 0.0)
    # mz = 0: gz = 1, so the perturbation is the strip profile fx; its peak
    # 4/sqrt(27) sin(s)(1 - cos(s)) = 1 is at s = 2 pi / 3, x = xb + (xe - xb)/3
    Xp = np.array([[xb - 1e-3, xb, xb + (xe - xb) / 3.0, xe, xe + 1e-3]]).T * np.ones((1, 2))
    p0 = cases.init_bc_wall_perturbation(1, xb, xe, 0, ZL, Xp, np.zeros_like(Xp), rng)
    assert np.all(p0[[0, 4]] == 0.0)
    assert abs(p0[1, 0]) < 1e-15 and abs(p0[3, 0]) < 1e-12
    assert abs(p0[2, 0] - 1.0) < 1e-13
    # mz < 0: a single spanwise mode sin(-2 pi mz z/ZL)
    p = cases.init_bc_wall_perturbation(1, xb, xe, -3, ZL, X, Z, rng)
    fx = cases.init_bc_wall_perturbation(1, xb, xe, 0, ZL, X, Z, rng)
    # (1e-13: the reference's PI literal 3.14159265358979 differs from pi by 3e-15)
    assert np.allclose(p, fx * np.sin(6.0 * math.pi * Z / ZL), rtol=0, atol=1e-13)
    # mz > 0: weights 1.25^-(m-1) normalised to 1 and one random phase per
    # mode, drawn first from the generator
    mz = 4
    ph = np.random.default_rng(7).random(mz)
    p = cases.init_bc_wall_perturbation(1, xb, xe, mz, ZL, X, Z, np.random.default_rng(7))
    w = 1.25 ** -np.arange(mz)
    w /= w.sum()
    gz = sum(w[m] * np.sin(2.0 * math.pi * (m + 1) * (Z / ZL + ph[m])) for m in range(mz))
    assert np.allclose(p, fx * gz, rtol=0, atol=1e-13)
    # type 2: uniform in (-1, 1) inside the strip, zero outside
    p2 = cases.init_bc_wall_perturbation(2, xb, xe, 0, ZL, X, Z, np.random.default_rng(3))
    inside = (X >= xb) & (X <= xe)
    assert np.all(p2[~inside] == 0.0) and np.all(np.abs(p2[inside]) < 1.0)
    assert np.std(p2[inside]) > 0.4


def test_temporal_factor(cases):
    assert cases.get_ht_multifrequancy(5, 0.0, np.random.default_rng(1))(0.37) == 1.0
    mt, beta = 3, 1000.0
    ph = np.random.default_rng(9).random(mt)
    ht = cases.get_ht_multifrequancy(mt, beta, np.random.default_rng(9))
    w = 1.25 ** -np.arange(mt)
    w /= w.sum()
    for tt in (0.0, 1.3e-4, 7.7e-3):
        ref = sum(w[m] * math.sin(2.0 * math.pi * ((m + 1) * beta * tt + ph[m])) for m in range(mt))
        assert abs(ht(tt) - ref) < 1e-13
    assert abs(ht(0.0) - sum(w[m] * math.sin(2 * math.pi * ph[m]) for m in range(mt))) < 1e-14
