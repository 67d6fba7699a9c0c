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


def curvilinear_box(n, a=0.05, L=2.0):
    """A periodic box of side L under the non-separable smooth map
    X = L (s_i + a sin 2pi s_j sin 2pi s_k), Y = L (s_j + a sin 2pi s_k sin 2pi s_i),
    Z = L (s_k + a sin 2pi s_i sin 2pi s_j), s = (index - 1)/n, given as 3-D
    coordinate arrays (the GRID3D input of init_mesh).  Returns the mesh, the
    span and the exact d(X, Y, Z)/d(xi, eta, zeta) with xi = (i - 1) h,
    h = 1/(n - 1) (OCFD_readpara.f90:351-353)."""
    load_package()
    from ocfd_amd import cases
    s = np.arange(n) / n
    S1, S2, S3 = np.meshgrid(s, s, s, indexing="ij")
    w = 2 * np.pi
    sn = lambda t: np.sin(w * t)
    cs = lambda t: np.cos(w * t)
    X = L * (S1 + a * sn(S2) * sn(S3))
    Y = L * (S2 + a * sn(S3) * sn(S1))
    Z = L * (S3 + a * sn(S1) * sn(S2))
    ds = (n - 1) / n            # d s / d xi
    Jm = np.empty((3, 3) + X.shape)   # Jm[c][e] = d coord c / d index-coordinate e
    Jm[0, 0] = L * ds * np.ones_like(X)
    Jm[0, 1] = L * a * w * cs(S2) * sn(S3) * ds
    Jm[0, 2] = L * a * w * sn(S2) * cs(S3) * ds
    Jm[1, 0] = L * a * w * sn(S3) * cs(S1) * ds
    Jm[1, 1] = L * ds * np.ones_like(X)
    Jm[1, 2] = L * a * w * cs(S3) * sn(S1) * ds
    Jm[2, 0] = L * a * w * cs(S1) * sn(S2) * ds
    Jm[2, 1] = L * a * w * sn(S1) * cs(S2) * ds
    Jm[2, 2] = L * ds * np.ones_like(X)
    mesh = []
    for A in (X, Y, Z):
        h = cases.halo_zeros(n, n, n, 1.0)
        cases.interior(h)[...] = A
        mesh.append(h)
    span = np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)
    return mesh, span, Jm


def bl_data_of_kinds(bc, ki, ku, nx, ny, nz, rng):
    """Inlet data of BcData_inlet = ki and upper data of BcData_upper = ku
    (OCFD_BC_Boundarylayers.f90:247-430 layouts) made from a deck's 1-D inlet
    profile and 0-D upper state, with small seeded variations along the face
    so that every position is distinguishable."""
    inl1, up0 = np.asarray(bc["inlet1d"]), np.asarray(bc["upper0d"])
    if ki == 0:
        inl = inl1[ny // 2].copy()
    elif ki == 1:
        inl = inl1.copy(order="F")
    else:
        inl = np.asfortranarray(inl1[:, None, :] * (1.0 + 0.01 * rng.uniform(-1, 1, (ny, nz, 1))))
    if ku == -1:
        up = None
    elif ku == 0:
        up = up0.copy()
    elif ku == 1:
        up = np.asfortranarray(up0[None, :] * (1.0 + 0.01 * rng.uniform(-1, 1, (nx, 1))))
    else:
        up = np.asfortranarray(up0[None, None, :] * (1.0 + 0.01 * rng.uniform(-1, 1, (nx, nz, 1))))
    return inl, up
