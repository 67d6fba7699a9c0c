"""The library's readers and writers of the reference's files, on the GPU.

* ocfd_read_flow_data (read_flow_data + init3d) of a restart written in the
  reference's layout gives the state that uploading the same primitives
  gives (bit for bit, after comput_flow_variables + OCFD_bc), for one
  species and for the 9-species real gas; ocfd_save_flow_data writes the
  resident primitives back in the same layout.
* ocfd_read_mesh of OCFD-grid.dat files of every Iflag_Gridtype gives the
  metrics ocfd_set_mesh gives for the same coordinates (bit for bit; the
  axisymmetric grid to rounding of cos/sin), and GRID_AND_JACOBIAN3D uses
  the metrics stored in the file.
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


def _prims(st, ns):
    I = fc.interior
    return [I(st[k]).copy(order="F") for k in ("D", "U", "V", "W", "T")] + [
        np.asfortranarray(np.stack([I(st["DI"][..., m]) for m in range(ns)], axis=3))]


@pytest.mark.parametrize("which", ["flat_plate", "jet"])
def test_restart_read_equals_upload(pkg, which, tmp_path):
    from ocfd_amd import formats
    case = fc.small_flat_plate(48, 32, 16) if which == "flat_plate" else fc.small_jet()
    cfg, mesh, span, st, bc, info = case
    ns = cfg.n_spec
    d, u, v, w, T, di = _prims(st, ns)
    rng = np.random.default_rng(5)
    u = u * (1.0 + 0.01 * rng.standard_normal(u.shape))
    path = str(tmp_path / "opencfd-comb.dat")
    formats.write_flow_data(path, 42, 1.25e-4, d, u, v, w, T, di)
    st2 = dict(st)
    st2["U"] = st["U"].copy(order="F")
    fc.interior(st2["U"])[...] = u
    out = []
    for mode in ("upload", "file"):
        g = pkg.FlowSolver(cfg)
        g.set_mesh(*mesh, span)
        if ns > 1:
            from ocfd_amd import cases
            cases.jet_gas(g, info["mech"])
            g.set_bc_jet(bc["inlet2d"])
        else:
            g.set_bc_boundarylayer(bc["inlet1d"], bc["upper0d"], bc["nonreflect_upper"], bc["wall_pertb"])
            g.set_bc_time(info["ht"](0.0))
        if mode == "upload":
            if ns == 1:
                from ocfd_amd import cases
                st2 = cases.state_from_primitives(d, u, v, w, T)
            else:
                E = info["mech"].energy(np.stack([di[..., m].ravel(order="F") for m in range(ns)]), T.ravel(order="F"))
                f = st2["F"].copy(order="F")
                f[..., 0], f[..., 1], f[..., 2], f[..., 3] = d, d * u, d * v, d * w
                f[..., 4] = E.reshape(d.shape, order="F") + d * (u * u + v * v + w * w) * 0.5
                st2["F"] = f
            g.upload_state(st2)
        else:
            assert g.read_flow_data(path) == (42, 1.25e-4)
        g.comput_flow_variables()
        g.OCFD_bc()
        out.append({n: g.download(n) for n in ("F", "D", "U", "V", "W", "T", "P", "ET", "DI")})
        if mode == "file":
            g.save_flow_data(str(tmp_path / "saved.dat"), 43, 2.5e-4)
            istep, tt, f = formats.read_flow_data(str(tmp_path / "saved.dat"), cfg.nx, cfg.ny, cfg.nz, ns)
            assert (istep, tt) == (43, 2.5e-4)
            for name, key in (("d", "D"), ("u", "U"), ("v", "V"), ("w", "W"), ("T", "T")):
                assert np.array_equal(f[name], fc.interior(out[-1][key])), name
            assert np.array_equal(f["di"], fc.interior(out[-1]["DI"]))
        g.close()
    a, b = out
    for n in a:
        x, y = (a[n], b[n]) if n == "F" else (fc.interior(a[n]), fc.interior(b[n]))
        if ns == 1:
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), n
        else:   # host (numpy) vs device energy of the real gas: a few ulps
            assert fc.compare(x, y)[0] <= 1e-13, n


def _metrics(g, pkg):
    return {n: g.download(n) for n in pkg.METRICS}


@pytest.mark.parametrize("gridtype", [10, 20, 30, 31])
def test_read_mesh_equals_set_mesh(pkg, gridtype, tmp_path):
    from ocfd_amd import formats
    if gridtype == 10:
        case = fc.small_flat_plate(48, 32, 16)
        cfg, mesh, span = case[0], case[1], case[2]
    else:
        from ocfd_amd import cases
        cfg = cases.isotropic(16)[0]
        mesh, span, _ = fc.curvilinear_box(16, a=0.05, L=2.0 * np.pi)
        if gridtype == 20:   # x, y independent of k; z(k) uniform
            X, Y, Z = (fc.interior(a) for a in mesh)
            mesh = list(mesh)
            for a in range(2):
                c = fc.interior(mesh[a])
                c[...] = c[:, :, :1]
            fc.interior(mesh[2])[...] = Z[:1, :1, :]
            span = span.copy()
            span[6], span[7] = 0.0, 0.0
    X, Y, Z = (fc.interior(a) for a in mesh)
    ref = pkg.FlowSolver(cfg)
    ref.set_mesh(*mesh, span)
    mref = _metrics(ref, pkg)
    ref.close()
    path = str(tmp_path / "OCFD-grid.dat")
    if gridtype == 10:
        formats.write_grid(path, 10, X[:, 0, 0], Y[0, :, 0], Z[0, 0, :])
    elif gridtype == 20:
        formats.write_grid(path, 20, X[:, :, 0], Y[:, :, 0], Z[0, 0, :])
    elif gridtype == 30:
        formats.write_grid(path, 30, X, Y, Z)
    else:
        # metrics in the file: the interior of set_mesh's, scaled so the
        # reader must use them and not recompute
        mi = [fc.interior(mref[n]) * 1.5 for n in ("AKX", "AKY", "AKZ", "AIX", "AIY", "AIZ", "ASX", "ASY", "ASZ",
                                                     "AJAC")]
        formats.write_grid(path, 31, X, Y, Z, *mi)
    g = pkg.FlowSolver(cfg)
    g.read_mesh(path, gridtype, span)
    mg = _metrics(g, pkg)
    g.close()
    if gridtype != 31:
        for n in pkg.METRICS:
            assert np.array_equal(mg[n].view(np.uint64), mref[n].view(np.uint64)), n
    else:
        for n in ("AKX", "AIY", "ASZ", "AJAC", "AKY"):
            assert np.array_equal(fc.interior(mg[n]), fc.interior(mref[n]) * 1.5), n
        # A*1 = A* / Ajac of the file's values
        assert np.array_equal(fc.interior(mg["AKX1"]), fc.interior(mg["AKX"]) / fc.interior(mg["AJAC"]))


def test_read_mesh_axisymmetric(pkg, tmp_path):
    """GRID2D_AXIAL_SYMM: y = R cos(seta), z = R sin(seta) from (x, R, seta)."""
    from ocfd_amd import formats, cases
    n = 16
    cfg = pkg.make_cfg(n, n, n, periodic=(0, 0, 0))
    x = np.linspace(0.0, 1.0, n)
    R = np.linspace(0.5, 1.5, n)
    seta = np.linspace(0.0, 0.6, n)
    X2, R2 = np.meshgrid(x, R, indexing="ij")
    path = str(tmp_path / "grid.dat")
    formats.write_grid(path, 21, X2, R2, seta)
    mesh = [cases.halo_zeros(n, n, n, 1.0) for _ in range(3)]
    fc.interior(mesh[0])[...] = X2[:, :, None]
    fc.interior(mesh[1])[...] = R2[:, :, None] * np.cos(seta)[None, None, :]
    fc.interior(mesh[2])[...] = R2[:, :, None] * np.sin(seta)[None, None, :]
    ref = pkg.FlowSolver(cfg)
    ref.set_mesh(*mesh)
    mref = _metrics(ref, pkg)
    ref.close()
    g = pkg.FlowSolver(cfg)
    g.read_mesh(path, 21)
    mg = _metrics(g, pkg)
    g.close()
    for nme in pkg.METRICS:
        d, _ = fc.compare(fc.interior(mg[nme]), fc.interior(mref[nme]))
        assert d <= 1e-12, (nme, d)
