"""GPU parity of filtering (OCFD_filtering.f90:4-362) through the C-ABI
against the oracle: the 9-point filter and the shock-capturing filter, each
direction and all three in turn, periodic and walled directions, boxes that
cut the domain, a second filter row on the same step, and the step interval.
Bar: bit for bit (the filters are sums and products in the reference's
order, FP64 without contraction on both sides)."""
import numpy as np
import pytest

from conftest import load_package, gpu_available
import flowcase as fc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_package()


def row(every=1, xyz=(1, 1, 1), box=(1, 999, 1, 999, 1, 999), scheme=1, s0=1.0, rth=1e-5):
    return [every, *xyz, *box, scheme, s0, rth]


CASES = {
    "fo9p_all": [row()],
    "fo9p_x_box": [row(xyz=(1, 0, 0), box=(3, 30, 2, 20, 1, 7), s0=0.6)],
    "fo9p_yz": [row(xyz=(0, 1, 1), s0=0.3)],
    "shock_all": [row(scheme=2, rth=1e-6)],
    "shock_y_box": [row(xyz=(0, 1, 0), box=(5, 40, 1, 999, 3, 9), scheme=2, s0=0.8, rth=1e-8)],
    "two_rows": [row(every=2, xyz=(1, 0, 1), s0=0.5), row(every=3, xyz=(1, 1, 0), scheme=2, rth=1e-7)],
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("which", ["flat_plate", "isotropic"])
def test_filtering_matches_oracle(pkg, name, which):
    import oracle
    case = fc.small_flat_plate(48, 32, 16) if which == "flat_plate" else fc.small_isotropic(16)
    g = pkg.FlowSolver(case[0])
    o = oracle.Oracle(pkg, case[0])
    fc.setup(g, case)
    fc.setup(o, case)
    dt1 = 0.5 * case[5]["dt"]
    for eng in (g, o):       # a developed state: two RK3 steps (one Istep)
        eng.flow_time_advance(dt1)
        eng.flow_time_advance(dt1)
    print(f"filtering: {which}, {name}", flush=True)
    for istep in (1, 2, 3, 6):
        g.filtering(istep, CASES[name])
        o.filtering(istep, CASES[name])
        g.synchronize()
        rep = fc.compare_state(g, o, names=("F",) + fc.PRIMS, tol=0.0)
        assert all(v[1] == 1.0 for v in rep.values()), (istep, rep)
    # the step after filtering reads the filtered f
    g.flow_time_advance(dt1)
    o.flow_time_advance(dt1)
    rep = fc.compare_state(g, o, tol=0.0)
    assert all(v[1] == 1.0 for v in rep.values()), rep
    g.close()
    o.close()


def test_filtering_argument_errors(pkg):
    case = fc.small_isotropic(16)
    g = pkg.FlowSolver(case[0])
    fc.setup(g, case)
    with pytest.raises(pkg.OcfdError):
        g.filtering(1, [row(every=0)])
    F = pkg.filter_array([row()])
    assert g.lib.ocfd_filtering(1, 11, pkg._ptr(F)) != 0
    assert g.lib.ocfd_filtering(1, 1, None) != 0
    f0 = g.download("F")
    g.filtering(1, [])                       # NFiltering = 0: nothing
    g.filtering(5, [row(every=2)])           # not this step: nothing
    assert g.download("F").tobytes() == f0.tobytes()
    g.close()
