"""GPU parity at the full flat-plate workload: 640 x 120 x 64 cells, hybrid
UD7L/WENO7/WENO5 with the type-2 sensor, Steger-Warming, CD6 viscous terms,
boundary-layer boundary conditions, periodic span.

The GPU runs the benchmarked configuration (certified fast sensor mean, the
viscous preparation and the step's start copy on a second stream, the
autotuner's tile shapes, the step replayed from a hipGraph); the oracle (the C
restatement, OpenMP over the host threads) runs the reference's loops.  Two
RK3 steps must agree bit for bit in the conserved variables and the
primitives, and so must the scheme masks of the last step.  The sensor field
itself is normalised by the certified tree mean, a few ulps from the
reference's sequential mean (<= 1e-12 relative).
"""
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


def test_flat_plate_640x120x64_two_steps_bitwise(pkg):
    import oracle
    from ocfd_amd import cases
    case = cases.flat_plate(640, 120, 64)
    cfg, info = case[0], case[5]
    dt1 = 0.5 * info["dt"]
    g = pkg.FlowSolver(cfg)
    fc.setup(g, case)
    g.set_option(pkg.OPT_HYBRID_SUM_MODE, 1)
    g.set_option(pkg.OPT_CONCURRENT, 1)
    print("tiles", g.autotune(2), flush=True)
    g.set_option(pkg.OPT_USE_GRAPH, 1)
    for _ in range(2):
        g.flow_time_advance(dt1)
    g.synchronize()
    got = {n: g.download(n) for n in ("F",) + fc.PRIMS}
    rh = fc.interior(g.download("RHYBRID"))
    masks = [g.download_mask(d) for d in range(3)]
    g.close()
    print("gpu done", flush=True)

    o = oracle.Oracle(pkg, cfg)
    print("oracle threads", o.set_threads(0), flush=True)
    fc.setup(o, case)
    for _ in range(2):
        o.flow_time_advance(dt1)
    for n, a in got.items():
        b = o.download(n)
        if n != "F":
            a, b = fc.interior(a), fc.interior(b)
        d, same = fc.compare(a, b, n)
        print(n, d, same, flush=True)
        assert np.all(np.isfinite(a)), n
        assert same == 1.0, (n, d, same)
    for d in range(3):
        om = o.download_mask(d)
        print("mask", d, np.bincount(om.ravel(), minlength=4), flush=True)
        assert np.array_equal(masks[d], om), d
    d, same = fc.compare(rh, fc.interior(o.download("RHYBRID")), "RHYBRID")
    print("RHYBRID", d, same, flush=True)
    assert d <= 1e-12
    o.close()
