"""GPU parity of the batched 1-D scheme entry (OCFD2d_flux1 / OCFD2d_flux2,
OCFD_inviscous.f90:209-362) against the oracle's restatement of the same
subroutines.  Integer-free floating point with -ffp-contract=off on both
sides: the bar is bit-identical hh(0:n) for every scheme, side, Bound_index
and Scheme_boundary combination."""
import ctypes
import itertools

import numpy as np
import pytest

from conftest import load_package, gpu_available

pytestmark = pytest.mark.gpu

LAP = 4
SCHEMES = {"WENO5": 1, "WENO7": 2, "OMP6": 3, "UD7L": 4, "HYBRID": 10}
DISS = 0.7  # UD7L_Diss: a blend of the upwind and central stencils


@pytest.fixture(scope="module")
def ctx():
    if not gpu_available():
        pytest.skip("no GPU")
    pkg = load_package()
    import oracle
    cfg = pkg.make_cfg(16, 16, 16, ud7l_diss=DISS)
    g = pkg.FlowSolver(cfg)
    o = oracle.Oracle(pkg, cfg)
    o.lib.orc_set_ud7l_diss(DISS)
    yield pkg, g, o
    g.close()


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def oracle_hh(o, scheme, side, v, scm, bidx, sb):
    npen, n = v.shape[0], v.shape[1] - 2 * LAP
    out = np.full((npen, n + 1), np.nan)
    b = np.array(bidx, dtype=np.int32)
    s = np.array(sb, dtype=np.int32)
    fn = o.lib.orc_recon_flux1 if side == 0 else o.lib.orc_recon_flux2
    for p in range(npen):
        vp = np.ascontiguousarray(v[p])
        hp = np.full(n + 1, np.nan)
        mp = np.ascontiguousarray(scm[p])
        fn(scheme, P(vp), P(hp), n, P(b), P(mp), P(s))
        out[p] = hp
    return out


def data(npen, n, seed):
    rng = np.random.default_rng(seed)
    # smooth pencils with a jump so the WENO weights leave their linear limit
    x = np.linspace(0.0, 1.0, n + 2 * LAP)
    v = np.sin(6.0 * x)[None, :] + 0.3 * rng.standard_normal((npen, n + 2 * LAP))
    v[:, (n + 2 * LAP) // 2:] += 2.0
    scm = rng.integers(1, 4, (npen, n + 1)).astype(np.uint8)
    return v, scm


BOUNDS = [((0, 0), (0, 0)), ((1, 1), (0, 0)), ((1, 0), (0, 0)), ((0, 1), (0, 0)),
          ((1, 1), (1, 1)), ((1, 1), (-1, -1)), ((1, 1), (0, 1)), ((1, 1), (-1, 0))]


@pytest.mark.parametrize("name", list(SCHEMES))
@pytest.mark.parametrize("side", [0, 1])
def test_flux1d_batch_bitwise(ctx, name, side):
    import torch
    pkg, g, o = ctx
    scheme = SCHEMES[name]
    for (n, npen, seed), (bidx, sb) in itertools.product([(9, 3, 1), (24, 5, 2), (37, 4, 3)], BOUNDS):
        v, scm = data(npen, n, seed)
        ref = oracle_hh(o, scheme, side, v, scm, bidx, sb)
        got = g.OCFD2d_flux(side, scheme, torch.from_numpy(v).cuda(), bidx, sb,
                            scm=torch.from_numpy(scm).cuda()).cpu().numpy()
        assert np.all(np.isfinite(ref)), (name, side, n, bidx, sb)
        same = np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        assert same, (name, side, n, bidx, sb, float(np.max(np.abs(got - ref))))


@pytest.mark.parametrize("name", list(SCHEMES))
@pytest.mark.parametrize("side", [0, 1])
def test_flux1d_batch_short_pencils(ctx, name, side):
    # n = 1, 2 and 3 with no boundary closure (Bound_index 0): every interface
    # reads the ghost points, as on a periodic or exchanged pencil
    import torch
    pkg, g, o = ctx
    for n, seed in ((1, 21), (2, 22), (3, 23)):
        v, scm = data(6, n, seed)
        ref = oracle_hh(o, SCHEMES[name], side, v, scm, (0, 0), (0, 0))
        got = g.OCFD2d_flux(side, SCHEMES[name], torch.from_numpy(v).cuda(), (0, 0), (0, 0),
                            scm=torch.from_numpy(scm).cuda()).cpu().numpy()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (name, side, n)


def test_flux1d_batch_large_batch_matches_pencils(ctx):
    # many pencils in one launch (more workgroups than CUs): every pencil equals
    # the same pencil computed alone, and the hybrid mask is honoured per pencil
    import torch
    pkg, g, o = ctx
    n, npen = 120, 3000
    v, scm = data(npen, n, 11)
    vt, mt = torch.from_numpy(v).cuda(), torch.from_numpy(scm).cuda()
    full = g.OCFD2d_flux(0, SCHEMES["HYBRID"], vt, (1, 1), (0, 0), scm=mt).cpu().numpy()
    for p in (0, 1, 1777, npen - 1):
        one = g.OCFD2d_flux(0, SCHEMES["HYBRID"], vt[p:p + 1], (1, 1), (0, 0), scm=mt[p:p + 1]).cpu().numpy()
        assert np.array_equal(full[p], one[0])
    pick = [0, 1777, npen - 1]
    ref = oracle_hh(o, SCHEMES["HYBRID"], 0, v[pick], scm[pick], (1, 1), (0, 0))
    assert np.array_equal(full[pick].view(np.uint64), ref.view(np.uint64))


def test_flux1d_batch_edges(ctx):
    import torch
    pkg, g, o = ctx
    # no pencils: nothing to do, no error
    out = g.OCFD2d_flux(0, SCHEMES["WENO7"], torch.empty((0, 16 + 2 * LAP), dtype=torch.float64, device="cuda"))
    assert tuple(out.shape) == (0, 17)
    v, scm = data(2, 16, 5)
    vt = torch.from_numpy(v).cuda()
    # the hybrid scheme without a mask is refused
    with pytest.raises(pkg.OcfdError):
        g.OCFD2d_flux(0, SCHEMES["HYBRID"], vt)
    # an unknown scheme and bad boundary flags are refused
    with pytest.raises(pkg.OcfdError):
        g.OCFD2d_flux(0, 5, vt)
    with pytest.raises(pkg.OcfdError):
        g.OCFD2d_flux(0, SCHEMES["WENO5"], vt, bound_index=(2, 0))
    with pytest.raises(pkg.OcfdError):
        g.OCFD2d_flux(2, SCHEMES["WENO5"], vt)
    # a hybrid mask value outside 1..3 gives NaN at that interior interface only
    scm[0, 8] = 0
    got = g.OCFD2d_flux(1, SCHEMES["HYBRID"], vt, scm=torch.from_numpy(scm).cuda()).cpu().numpy()
    assert np.isnan(got[0, 8]) and np.isfinite(np.delete(got[0], 8)).all() and np.isfinite(got[1]).all()
