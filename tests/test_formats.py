"""The reference's file formats on the host (CPU only).

* Control decks: the reference's own example decks (flat plate, jet,
  isotropic; copied to tests/golden/ by copy_reference_data.py) parse to the
  values they hold, absent keys take read_control_parameter's defaults
  (OCFD_readpara.f90:65-110), and the deck maps onto an ocfd_cfg.
* Fortran unformatted records: the flat-plate example's OCFD-grid.dat (a
  GRID1D file the reference ships) reads as three records of the deck's
  sizes; a restart written by a Fortran program (flang, the same statements
  as save_flow_data) reads back through formats.read_flow_data, and a file
  written by formats.write_flow_data reads back in Fortran with the values
  intact.
* Boundary-layer data files: the example's 1-D inflow and 0-D upper data,
  with the reference's rescaling of the partial densities.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import load_package, ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.fixture(scope="module")
def fm():
    load_package()
    from ocfd_amd import formats
    return formats


def test_flat_plate_deck(fm):
    nml = fm.read_namelist(os.path.join(GOLDEN, "flat225_opencfd-comb2.in"))
    assert (nml["nx_global"], nml["ny_global"], nml["nz_global"]) == (2193, 72, 64)
    assert (nml["npx0"], nml["npy0"], nml["npz0"], nml["n_spec"]) == (4, 1, 2, 2)
    assert nml["iflag_gridtype"] == fm.GRID1D and nml["scheme_invis"] == "OMP6" and nml["iperiodic_z"] == 1
    assert nml["periodic_kspan"] == [0.0, 0.0, 0.004445]
    assert nml["bc_para"][:14] == [1, 0, 0, 0, 321.936, 0.0, 1, 23.486, 0.1143, 0.1270, 5, 10, 0.004445, 28861.58]
    assert nml["bc_para"][14] == 0.0 and len(nml["bc_para"]) == 100
    assert nml["dt"] == 8.652e-9 and nml["istep_save"] == 10000
    assert nml["ana_para(:,1)"] == [100, 10, 10000]
    # defaults of read_control_parameter for keys the deck leaves out
    assert nml["pr"] == 0.77 and nml["flux_splitting"] == 1 and nml["ud7l_diss"] == 1.0
    assert nml["hybrid_para"][:4] == [1.0] * 4 and nml["scheme_boundary"] == [0] * 6
    cfg, run = fm.config_from_namelist(nml)
    m = load_package()
    assert (cfg.nx, cfg.ny, cfg.nz, cfg.n_spec) == (2193, 72, 64, 2)
    assert cfg.scheme_invis == m.SCHEME_OMP6 and cfg.scheme_vis == m.SCHEME_CD6
    assert tuple(cfg.periodic) == (0, 0, 1) and cfg.ibc == m.BC_BOUNDARYLAYER
    assert cfg.bc_para[4] == 321.936 and cfg.bc_para[7] == 23.486 and cfg.pr == 0.77
    assert run["span"][8] == 0.004445 and run["gridtype"] == 10 and run["dt"] == 8.652e-9


def test_jet_and_isotropic_decks(fm):
    m = load_package()
    jet = fm.read_namelist(os.path.join(GOLDEN, "jet_opencfd-comb2.in"))
    assert jet["n_spec"] == 9 and jet["if_reaction"] == 1 and jet["chemical_sub_step"] == 5
    assert (jet["flag_viscosity"], jet["flag_conductivity"], jet["flag_diffusion"]) == (1, 1, 1)
    assert jet["boundarycondition"] == "BC_JET" and jet["bc_para"][0] == 1
    cfg, run = fm.config_from_namelist(jet)
    assert cfg.ibc == m.BC_JET and cfg.if_real_gas == 1 and run["chemical_sub_step"] == 5
    iso = fm.read_namelist(os.path.join(GOLDEN, "isotropic_opencfd-comb2.in"))
    assert iso["scheme_invis"] == "UD7L" and iso["ud7l_diss"] == 0.5
    assert iso["periodic_ispan"] == [6.2831853071795864e-3, 0.0, 0.0]
    cfg, run = fm.config_from_namelist(iso)
    assert tuple(cfg.periodic) == (1, 1, 1) and cfg.ud7l_diss == 0.5 and cfg.ibc == m.BC_NONE


def test_namelist_syntax(fm):
    text = """&control_opencfd
      nx_global=8, ny_global = 9 nz_global=10  ! comment
      Scheme_Invis='HYBRID'  Hybrid_para(2)=3.5d0, 12.
      BC_Para = 3*2.0, 7 Ghost_Cell(3)=2
    /"""
    nml = fm.read_namelist(text)
    assert (nml["nx_global"], nml["ny_global"], nml["nz_global"]) == (8, 9, 10)
    assert nml["hybrid_para"][:4] == [1.0, 3.5, 12.0, 1.0]
    assert nml["bc_para"][:5] == [2.0, 2.0, 2.0, 7, 0.0]
    assert nml["ghost_cell"] == [0, 0, 2, 0, 0, 0]


def test_reference_grid_file(fm, tmp_path):
    """The GRID1D OCFD-grid.dat of the flat-plate example: three records
    x(2193), y(72), z(64) (read_mesh1d, OCFD_mesh.f90:221-250)."""
    path = os.path.join(GOLDEN, "flat225_OCFD-grid.dat")
    recs = fm.read_records(path)
    assert [len(r) for r in recs] == [8 * 2193, 8 * 72, 8 * 64]
    x, y, z = fm.read_grid(path, fm.GRID1D, 2193, 72, 64)
    assert np.all(np.diff(x) > 0) and np.all(np.diff(y) > 0) and y[0] == 0.0
    dz = np.diff(z)
    assert np.allclose(dz, 0.004445 / 64, rtol=1e-6)          # the deck's Periodic_KSpan over 64 planes
    assert np.diff(y)[-1] / np.diff(y)[0] > 10.0                # wall-normal stretching
    out = str(tmp_path / "grid.dat")
    fm.write_grid(out, fm.GRID1D, x, y, z)
    assert open(out, "rb").read() == open(path, "rb").read()


def test_boundary_layer_data_files(fm):
    """flow1d-inlet-comb.dat (72 rows of y, d, u, v, w, T, d1, d2) and
    flow0d-upper-comb.dat, partial densities rescaled to sum to d
    (init_bcdata_boundary, OCFD_BC_Boundarylayers.f90:276-296, 354-373)."""
    inl = fm.read_bc_data1d(os.path.join(GOLDEN, "flat225_flow1d-inlet-comb.dat"), 72, 2)
    assert inl.shape == (72, 7)
    assert inl[0, 4] == 321.936 and inl[0, 1] == 0.0               # isothermal, no-slip wall row
    assert np.allclose(inl[:, 5] + inl[:, 6], inl[:, 0], rtol=1e-15)
    raw = np.loadtxt(os.path.join(GOLDEN, "flat225_flow1d-inlet-comb.dat"), skiprows=1)
    assert np.array_equal(inl[:, :5], raw[:, 1:6])
    assert np.allclose(inl[:, 5] / inl[:, 6], raw[:, 6] / raw[:, 7], rtol=1e-14)
    up = fm.read_bc_data0d(os.path.join(GOLDEN, "flat225_flow0d-upper-comb.dat"), 2)
    assert up[4] == 169.66210155 and abs(up[5] + up[6] - up[0]) <= 1e-15 * up[0]


F_WRITE = """
program wr
  implicit none
  integer, parameter :: nx = 5, ny = 4, nz = 3, ns = 2
  real(8) :: a(nx, ny)
  integer :: i, j, k, m, istep
  real(8) :: tt
  istep = 1234; tt = 0.5d-3
  open(155, file='restart.dat', form='unformatted')
  write(155) istep, tt
  do m = 1, 5 + ns
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          a(i, j) = m * 1000 + i + 10 * j + 100 * k + 0.25d0
        end do
      end do
      write(155) a
    end do
  end do
  close(155)
end program
"""

F_READ = """
program rd
  implicit none
  integer, parameter :: nx = 5, ny = 4, nz = 3, ns = 2
  real(8) :: a(nx, ny), s
  integer :: k, m, istep
  real(8) :: tt
  open(55, file='py.dat', form='unformatted')
  read(55) istep, tt
  s = 0
  do m = 1, 5 + ns
    do k = 1, nz
      read(55) a
      s = s + sum(a * m)
    end do
  end do
  close(55)
  print '(I0,1X,ES24.16,1X,ES24.16)', istep, tt, s
end program
"""


def test_restart_format_against_fortran(fm, tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler")
    (tmp_path / "wr.f90").write_text(F_WRITE)
    (tmp_path / "rd.f90").write_text(F_READ)
    for p in ("wr", "rd"):
        subprocess.run([FLANG, f"{p}.f90", "-o", p], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run(["./wr"], cwd=tmp_path, check=True)
    istep, tt, f = fm.read_flow_data(str(tmp_path / "restart.dat"), 5, 4, 3, 2)
    assert istep == 1234 and tt == 0.5e-3
    I, J, K = np.meshgrid(np.arange(1, 6), np.arange(1, 5), np.arange(1, 4), indexing="ij")
    base = I + 10 * J + 100 * K + 0.25
    for m, name in enumerate(("d", "u", "v", "w", "T"), 1):
        assert np.array_equal(f[name], m * 1000 + base), name
    for s in range(2):
        assert np.array_equal(f["di"][..., s], (6 + s) * 1000 + base)
    fm.write_flow_data(str(tmp_path / "py.dat"), 77, 2.5, f["d"], f["u"], f["v"], f["w"], f["T"], f["di"])
    out = subprocess.run(["./rd"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    expect = sum(m * float(np.sum(m * 1000 + base)) for m in range(1, 8))
    assert int(out[0]) == 77 and float(out[1]) == 2.5 and abs(float(out[2]) - expect) <= 1e-12 * expect
    assert open(tmp_path / "py.dat", "rb").read() == open(tmp_path / "restart.dat", "rb").read().replace(
        np.array([1234], "<i4").tobytes() + np.array([0.5e-3], "<f8").tobytes(),
        np.array([77], "<i4").tobytes() + np.array([2.5], "<f8").tobytes(), 1)
