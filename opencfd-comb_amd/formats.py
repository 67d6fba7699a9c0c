"""The reference's input and output files on the host side.

* The control deck ``opencfd-comb2.in``: the Fortran namelist
  ``control_opencfd`` read by read_control_parameter
  (code/mpi-io/OCFD_readpara.f90:41-110), with the reference's defaults for
  absent keys, and its translation into an ``ocfd_cfg``
  (set_parameters / set_scheme_para, OCFD_readpara.f90:347-431).
* Fortran unformatted sequential records (a 4-byte length before and after
  each record): the restart files ``opencfd-comb.dat`` / ``OCFD-comb%08d.dat``
  (OCFD_IO.f90:5-64), the grid ``OCFD-grid.dat`` (OCFD_mesh.f90:187-317) and
  the 2-D boundary data ``flow2d-inlet-comb.dat`` / ``flow2d-upper-comb.dat``.
* The boundary-layer text data ``flow0d-*-comb.dat`` / ``flow1d-*-comb.dat``
  (init_bcdata_boundary, OCFD_BC_Boundarylayers.f90:246-400), with the
  reference's rescaling of the partial densities to the density.

The device-side readers of the restart and the grid are the library's
(``FlowSolver.read_flow_data`` / ``save_flow_data`` / ``read_mesh``); the
functions here serve decks, tests and host-side tools.
"""
import re

import numpy as np

from . import (make_cfg, SCHEME_WENO5, SCHEME_WENO7, SCHEME_OMP6, SCHEME_UD7L, SCHEME_CD6, SCHEME_CD8,
               SCHEME_HYBRID, BC_NONE, BC_BOUNDARYLAYER, BC_JET)

GRID1D, GRID2D_PLANE, GRID2D_AXIAL_SYMM, GRID3D, GRID_AND_JACOBIAN3D = 10, 20, 21, 30, 31

# scheme names of set_scheme_para (OCFD_readpara.f90:378-405) and BC names (:312-331)
SCHEMES = {"WENO5": SCHEME_WENO5, "WENO7": SCHEME_WENO7, "OMP6": SCHEME_OMP6, "UD7L": SCHEME_UD7L,
           "HYBRID": SCHEME_HYBRID, "CD6": SCHEME_CD6, "CD8": SCHEME_CD8}
BCS = {"BC_NONE": BC_NONE, "BC_BOUNDARYLAYER": BC_BOUNDARYLAYER, "BC_JET": BC_JET}


# ---------------------------------------------------------------------------
# namelist
# ---------------------------------------------------------------------------
def namelist_defaults():
    """The defaults read_control_parameter sets before reading the deck
    (OCFD_readpara.f90:65-110)."""
    return dict(npx0=1, npy0=1, npz0=1, lap=4, iperiodic_x=0, iperiodic_y=0, iperiodic_z=0, scheme_invis="WENO5",
                scheme_vis="CD6", if_scheme_character=0, dt=1e-9, end_time=100.0, iflag_gridtype=GRID3D,
                periodic_ispan=[0.0] * 3, periodic_jspan=[0.0] * 3, periodic_kspan=[0.0] * 3, bc_para=[0.0] * 100,
                boundarycondition="BC_None", istep_show=1, istep_save=1000, if_viscous=1, scheme_boundary=[0] * 6,
                if_mass_force=0, mass_force=[0.0] * 3, aoa=0.0, ana_number=0, nfiltering=0, ud7l_diss=1.0,
                hybrid_para=[1.0] * 100, flux_splitting=1, ghost_cell=[0] * 6, flag_viscosity=0,
                flag_conductivity=0, flag_diffusion=0, amu0=1.789e-5, pr=0.77, sc=1.0, chemical_sub_step=1,
                if_reaction=0, if_realgas=1, chem_timeadv=1, if_debug=0, iflag_particle=0, n_spec=1)


def _value(tok):
    t = tok.strip()
    if not t:
        return None
    if t[0] in "'\"":
        return t.strip("'\"")
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    if re.fullmatch(r"[+-]?\d+", t):
        return int(t)
    return float(low.replace("d", "e"))


def _split_values(text):
    """Comma/blank separated list-directed values, with r*value repeats."""
    out = []
    for tok in re.findall(r"'[^']*'|\"[^\"]*\"|[^,\s]+", text):
        m = re.fullmatch(r"(\d+)\*(.+)", tok)
        if m:
            out.extend([_value(m.group(2))] * int(m.group(1)))
        else:
            out.append(_value(tok))
    return out


def read_namelist(text_or_path, group="control_opencfd"):
    """The namelist group `group` of a deck (path or text) as a dict with
    lower-case keys, over the reference's defaults.  Array elements given as
    KEY(i) or KEY(:,j) set those elements (1-based); KEY = v1, v2, ... sets
    the leading elements.  '!' starts a comment; '&' or '$' open and
    '/', '&end' or '$end' close the group."""
    text = text_or_path
    if "\n" not in text_or_path:
        text = open(text_or_path).read()
    lines = [re.sub(r"!.*", "", ln) for ln in text.splitlines()]
    body = " ".join(lines)
    m = re.search(r"[&$]" + group + r"\b(.*?)(?:[&$]end\b|/)", body, re.IGNORECASE | re.DOTALL)
    if not m:
        raise ValueError(f"namelist group {group} not found")
    nml = namelist_defaults()
    parts = re.split(r"([A-Za-z_][A-Za-z_0-9]*(?:\([^)]*\))?)\s*=", m.group(1))
    for k in range(1, len(parts) - 1, 2):
        key, vals = parts[k].strip().lower(), _split_values(parts[k + 1])
        name, _, sub = key.partition("(")
        if not sub:
            cur = nml.get(name)
            if isinstance(cur, list) and (len(vals) != 1 or len(cur) > 1):
                cur = list(cur)
                cur[:len(vals)] = vals
                nml[name] = cur
            else:
                nml[name] = vals[0] if len(vals) == 1 else vals
        else:
            idx = sub.rstrip(")").split(",")
            cur = nml.get(name)
            if isinstance(idx[0], str) and idx[0].strip() in (":", ""):
                col = int(idx[1]) if len(idx) > 1 else 1
                nml[f"{name}(:,{col})"] = vals
            else:
                i0 = int(idx[0]) - 1
                if not isinstance(cur, list):
                    cur = [] if cur is None else [cur]
                cur = list(cur) + [None] * max(0, i0 + len(vals) - len(cur))
                cur[i0:i0 + len(vals)] = vals
                nml[name] = cur
    return nml


def config_from_namelist(nml):
    """(ocfd_cfg, run parameters) from a parsed deck: the fields the hot path
    needs (set_parameters, OCFD_readpara.f90:347-431) and the driver's own
    (dt, End_time, Istep_show/save, Iflag_Gridtype, spans, chemistry)."""
    scheme_invis = SCHEMES[str(nml["scheme_invis"]).upper()]
    scheme_vis = SCHEMES[str(nml["scheme_vis"]).upper()]
    ibc = BCS.get(str(nml["boundarycondition"]).upper())
    if ibc is None:
        raise ValueError(f"boundary condition {nml['boundarycondition']} is not on this path")
    bc = [float(x) for x in nml["bc_para"][:32]]
    cfg = make_cfg(int(nml["nx_global"]), int(nml["ny_global"]), int(nml["nz_global"]),
                   periodic=(nml["iperiodic_x"], nml["iperiodic_y"], nml["iperiodic_z"]),
                   scheme_invis=scheme_invis, scheme_vis=scheme_vis, flux_splitting=int(nml["flux_splitting"]),
                   if_viscous=int(nml["if_viscous"]), ibc=ibc, ud7l_diss=float(nml["ud7l_diss"]),
                   hybrid_para=[float(x) for x in nml["hybrid_para"][:16]], amu0=float(nml["amu0"]),
                   pr=float(nml["pr"]), sc=float(nml["sc"]),
                   mass_force=[float(x) for x in nml["mass_force"]] if int(nml["if_mass_force"]) else None,
                   bc_para=bc, scheme_boundary=[int(x) for x in nml["scheme_boundary"]],
                   n_spec=int(nml["n_spec"]), if_real_gas=int(nml["if_realgas"]) if int(nml["n_spec"]) > 1 else 0,
                   flag_vis=int(nml["flag_viscosity"]), flag_cond=int(nml["flag_conductivity"]),
                   flag_diff=int(nml["flag_diffusion"]), if_scheme_character=int(nml["if_scheme_character"]))
    for a in range(6):
        cfg.ghost_cell[a] = int(nml["ghost_cell"][a])
    span = np.array(list(nml["periodic_ispan"]) + list(nml["periodic_jspan"]) + list(nml["periodic_kspan"]),
                    dtype=np.float64)
    run = dict(dt=float(nml["dt"]), end_time=float(nml["end_time"]), istep_show=int(nml["istep_show"]),
               istep_save=int(nml["istep_save"]), gridtype=int(nml["iflag_gridtype"]), span=span,
               chemical_sub_step=int(nml["chemical_sub_step"]), chem_timeadv=int(nml["chem_timeadv"]),
               if_reaction=int(nml["if_reaction"]), nfiltering=int(nml["nfiltering"]),
               npx0=int(nml["npx0"]), npy0=int(nml["npy0"]), npz0=int(nml["npz0"]))
    return cfg, run


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records
# ---------------------------------------------------------------------------
def read_records(path):
    """Every record of a Fortran unformatted sequential file, as bytes."""
    data = open(path, "rb").read()
    out, p = [], 0
    while p < len(data):
        n = int(np.frombuffer(data, dtype="<i4", count=1, offset=p)[0])
        if n < 0 or p + 8 + n > len(data):
            raise ValueError(f"{path}: bad record marker at byte {p}")
        out.append(data[p + 4:p + 4 + n])
        if int(np.frombuffer(data, dtype="<i4", count=1, offset=p + 4 + n)[0]) != n:
            raise ValueError(f"{path}: record markers differ at byte {p}")
        p += 8 + n
    return out


def write_records(path, records):
    with open(path, "wb") as f:
        for r in records:
            b = bytes(r)
            n = np.array([len(b)], dtype="<i4").tobytes()
            f.write(n + b + n)


def _field3d(recs, pos, nx, ny, nz):
    a = np.stack([np.frombuffer(recs[pos + k], dtype="<f8").reshape((nx, ny), order="F") for k in range(nz)],
                 axis=2)
    return np.asfortranarray(a), pos + nz


def read_flow_data(path, nx, ny, nz, n_spec):
    """(Istep, tt, dict of (nx, ny, nz) arrays d, u, v, w, T, di[..., n_spec])."""
    recs = read_records(path)
    istep = int(np.frombuffer(recs[0], dtype="<i4", count=1)[0])
    tt = float(np.frombuffer(recs[0], dtype="<f8", count=1, offset=4)[0])
    out, pos = {}, 1
    for name in ("d", "u", "v", "w", "T"):
        out[name], pos = _field3d(recs, pos, nx, ny, nz)
    di = []
    for _ in range(n_spec):
        a, pos = _field3d(recs, pos, nx, ny, nz)
        di.append(a)
    out["di"] = np.asfortranarray(np.stack(di, axis=3))
    return istep, tt, out


def write_flow_data(path, istep, tt, d, u, v, w, T, di):
    """save_flow_data's layout from (nx, ny, nz) arrays and di (nx, ny, nz, n_spec)."""
    head = np.array([istep], dtype="<i4").tobytes() + np.array([tt], dtype="<f8").tobytes()
    recs = [head]
    fields = [d, u, v, w, T] + [di[..., m] for m in range(di.shape[3])]
    for a in fields:
        a = np.asarray(a, dtype="<f8")
        recs.extend(np.asfortranarray(a[:, :, k]).tobytes(order="F") for k in range(a.shape[2]))
    write_records(path, recs)


def read_grid(path, gridtype, nx, ny, nz):
    """OCFD-grid.dat: GRID1D -> (x(nx), y(ny), z(nz)); GRID2D_PLANE /
    GRID2D_AXIAL_SYMM -> (a(nx,ny), b(nx,ny), c(nz)); GRID3D -> (X, Y, Z) of
    shape (nx, ny, nz); GRID_AND_JACOBIAN3D -> those and the ten metric
    arrays."""
    recs = read_records(path)
    f8 = lambda r: np.frombuffer(r, dtype="<f8")
    if gridtype == GRID1D:
        return tuple(f8(recs[a]).copy() for a in range(3))
    if gridtype in (GRID2D_PLANE, GRID2D_AXIAL_SYMM):
        return (f8(recs[0]).reshape((nx, ny), order="F").copy(), f8(recs[1]).reshape((nx, ny), order="F").copy(),
                f8(recs[2]).copy())
    out, pos = [], 0
    for _ in range(3 if gridtype == GRID3D else 13):
        a, pos = _field3d(recs, pos, nx, ny, nz)
        out.append(a)
    return tuple(out)


def write_grid(path, gridtype, *arrays):
    """Inverse of read_grid."""
    recs = []
    for a in arrays:
        a = np.asarray(a, dtype="<f8")
        if gridtype in (GRID3D, GRID_AND_JACOBIAN3D):
            recs.extend(np.asfortranarray(a[:, :, k]).tobytes(order="F") for k in range(a.shape[2]))
        else:
            recs.append(np.asfortranarray(a).tobytes(order="F"))
    write_records(path, recs)


# ---------------------------------------------------------------------------
# boundary-layer data (init_bcdata_boundary)
# ---------------------------------------------------------------------------
def _rescale(rows, n_spec):
    """di(:) scaled so that they sum to d, per row (OCFD_BC_Boundarylayers.f90:285-295)."""
    rows = np.array(rows, dtype=np.float64)
    d0 = np.zeros(rows.shape[:-1])
    for m in range(n_spec):
        d0 = d0 + rows[..., 5 + m]
    for m in range(n_spec):
        rows[..., 5 + m] = rows[..., 5 + m] * rows[..., 0] / d0
    return rows


def read_bc_data0d(path, n_spec):
    """flow0d-inlet/upper-comb.dat: two header lines, then d, u, v, w, T, di(:)."""
    toks = " ".join(open(path).read().splitlines()[2:]).split()
    return _rescale([float(t.replace("d", "e").replace("D", "e")) for t in toks[:5 + n_spec]], n_spec)


def read_bc_data1d(path, n, n_spec):
    """flow1d-inlet/upper-comb.dat: a header line, then n lines of
    (coordinate, d, u, v, w, T, di(:)); returns (n, 5 + n_spec)."""
    lines = [ln for ln in open(path).read().splitlines()[1:] if ln.strip()]
    rows = [[float(t.replace("d", "e").replace("D", "e")) for t in ln.split()[1:6 + n_spec]] for ln in lines[:n]]
    if len(rows) != n:
        raise ValueError(f"{path}: {len(rows)} rows, {n} expected")
    return _rescale(rows, n_spec)


def read_bc_data2d(path, n1, n2, n_spec):
    """flow2d-inlet/upper-comb.dat: 5 + n_spec unformatted records of
    (n1, n2) reals; returns (n1, n2, 5 + n_spec)."""
    recs = read_records(path)
    out = np.stack([np.frombuffer(recs[m], dtype="<f8").reshape((n1, n2), order="F") for m in range(5 + n_spec)],
                   axis=2)
    return np.asfortranarray(_rescale(out, n_spec))
