"""Shared set-up for flow-path tests: build a case, load it into the HIP
solver and/or the oracle, compare fields."""
import numpy as np

from conftest import load_package

PRIMS = ("D", "U", "V", "W", "T", "P", "ET")


def small_isotropic(n=16, **kw):
    m = load_package()
    from ocfd_amd import cases
    cfg, mesh, span, st, info = cases.isotropic(n, **kw)
    return cfg, mesh, span, st, None, info


def small_flat_plate(nx=48, ny=32, nz=16, **kw):
    m = load_package()
    from ocfd_amd import cases
    cfg, mesh, span, st, bc, info = cases.flat_plate(nx, ny, nz, lx=0.2 * nx / 640.0, ly=0.012, **kw)
    return cfg, mesh, span, st, bc, info


def setup(engine, case, ht=None):
    cfg, mesh, span, st, bc, info = case
    engine.set_mesh(*mesh, span)
    if bc is not None:
        engine.set_bc_boundarylayer(bc["inlet1d"], bc["upper0d"], bc["nonreflect_upper"], bc["wall_pertb"])
        engine.set_bc_time(info["ht"](0.0) if ht is None else ht)
    engine.upload_state(st)
    # NS_solver prologue: comput_flow_variables + OCFD_bc (OCFD_NS_Solver.f90:16-17)
    engine.comput_flow_variables()
    engine.OCFD_bc()
    return engine


def interior(a, lap=4):
    return a[lap:-lap, lap:-lap, lap:-lap]


def compare(a, b, name=""):
    """Return (max relative difference scaled by max|b|, fraction bit-identical)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    diff = float(np.max(np.abs(a - b))) / scale
    same = float(np.mean(a.view(np.uint64) == b.view(np.uint64))) if a.dtype == np.float64 else float(np.mean(a == b))
    return diff, same


def compare_state(gpu, orc, names=("F",) + PRIMS, tol=1e-12):
    """Compare interior values of the named fields; return a report dict and
    assert every conserved variable / primitive is within tol (relative to its
    max magnitude)."""
    rep = {}
    for name in names:
        g, o = gpu.download(name), orc.download(name)
        if name in PRIMS:
            g, o = interior(g), interior(o)
        if name == "F":
            for mvar in range(5):
                rep[f"F{mvar + 1}"] = compare(g[..., mvar], o[..., mvar], name)
        else:
            rep[name] = compare(g, o, name)
    bad = {k: v for k, v in rep.items() if not (v[0] <= tol)}
    assert not bad, f"parity failure (max rel diff, bit-identical fraction): {bad}"
    return rep


def small_jet(nx=32, ny=24, nz=24, **kw):
    load_package()
    from ocfd_amd import cases
    return cases.jet(nx, ny, nz, **kw)


def setup_jet(engine, case):
    """Species thermo/transport, jet inflow data, state, then the NS_solver
    prologue (comput_flow_variables + OCFD_bc)."""
    from ocfd_amd import cases
    cfg, mesh, span, st, bc, info = case
    engine.set_mesh(*mesh, span)
    cases.jet_gas(engine, info["mech"])
    if bc is not None and "inlet2d" in bc:
        engine.set_bc_jet(bc["inlet2d"])
    engine.upload_state(st)
    engine.comput_flow_variables()
    engine.OCFD_bc()
    return engine


SPECIES_PRIMS = PRIMS + ("DI",)
