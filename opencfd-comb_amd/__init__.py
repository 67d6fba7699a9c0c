"""MI355X-native per-step hot path of OpenCFD-Comb (host side).

The compute lives in ``libocfd_hip.so`` (HIP, gfx950), built from ``csrc/`` by
``build.py`` and reached through the C-ABI declared in
``include/ocfd_hip.h``.  This module is the Python mirror of the reference's
subroutine interface for the path: ``FlowSolver`` exposes
``flow_time_advance(dt1)`` and the per-stage pieces under the reference's
names (``code/flow-solver/OCFD_Time_Adv.f90:7-130``,
``code/flow-solver/OCFD_inviscous.f90:47``, ``code/flow-solver/OCFD_viscous.f90:27``,
``code/boundary-condition/OCFD_BC.f90:2``).

Arrays follow the reference's Fortran layout: halo arrays are numpy arrays of
shape ``(nx+8, ny+8, nz+8)`` in Fortran order (index 0 is the reference's
``1-LAP``), interior 5-variable arrays have shape ``(nx, ny, nz, 5)``.

There is no CPU fallback: if the HIP library is missing every entry point
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libocfd_hip.so")
LAP = 4

# codes of OCFD_constants (code/head/MODULE_parameter.f90:14-26)
SCHEME_WENO5, SCHEME_WENO7, SCHEME_OMP6, SCHEME_UD7L, SCHEME_CD6, SCHEME_CD8 = 1, 2, 3, 4, 5, 6
SCHEME_HYBRID = 10
SPLIT_SW, SPLIT_LLF = 1, 2
BC_NONE, BC_BOUNDARYLAYER, BC_JET = 0, 2, 4

# field identifiers (include/ocfd_hip.h)
FIELDS = ["F", "FN", "DU", "DI", "D", "U", "V", "W", "T", "P", "ET", "AMU", "AMK", "RHYBRID",
          "AKX", "AKY", "AKZ", "AIX", "AIY", "AIZ", "ASX", "ASY", "ASZ", "AJAC",
          "AKX1", "AKY1", "AKZ1", "AIX1", "AIY1", "AIZ1", "ASX1", "ASY1", "ASZ1", "DIN",
          "AMD", "DUS", "QP"]
FIELD = {name: i for i, name in enumerate(FIELDS)}
METRICS = FIELDS[FIELD["AKX"]:FIELD["ASZ1"] + 1]

OPT_HYBRID_SUM_MODE, OPT_USE_GRAPH, OPT_KERNEL_TIMING, OPT_COLUMN_PLANES, OPT_CONCURRENT = 1, 2, 3, 4, 5
OPT_HYBRID_CERT_WIDEN = 6
OPT_METRIC_TABLES = 7
CHEM_FIXED, CHEM_ADAPTIVE, CHEM_SURROGATE = 0, 1, 2


class OcfdCfg(ctypes.Structure):
    """Mirror of ``ocfd_cfg`` (include/ocfd_hip.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lap", ctypes.c_int32), ("n_spec", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("scheme_invis", ctypes.c_int32), ("scheme_vis", ctypes.c_int32),
        ("flux_splitting", ctypes.c_int32), ("if_viscous", ctypes.c_int32),
        ("if_scheme_character", ctypes.c_int32),
        ("scheme_boundary", ctypes.c_int32 * 6), ("ghost_cell", ctypes.c_int32 * 6),
        ("ibc", ctypes.c_int32), ("if_mass_force", ctypes.c_int32), ("if_real_gas", ctypes.c_int32),
        ("flag_vis", ctypes.c_int32), ("flag_cond", ctypes.c_int32), ("flag_diff", ctypes.c_int32),
        ("wall_catalysis", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("ud7l_diss", ctypes.c_double), ("hybrid_para", ctypes.c_double * 16),
        ("amu0", ctypes.c_double), ("pr", ctypes.c_double), ("sc", ctypes.c_double),
        ("mass_force", ctypes.c_double * 3), ("bc_para", ctypes.c_double * 32),
    ]


def make_cfg(nx, ny, nz, *, periodic=(1, 1, 1), scheme_invis=SCHEME_WENO7, scheme_vis=SCHEME_CD6,
             flux_splitting=SPLIT_SW, if_viscous=1, ibc=BC_NONE, ud7l_diss=1.0, hybrid_para=None,
             amu0=1.789e-5, pr=0.72, sc=1.0, mass_force=None, bc_para=None, scheme_boundary=None,
             n_spec=1, if_real_gas=0, flag_vis=0, flag_cond=0, flag_diff=0, if_scheme_character=0):
    """Build an ``ocfd_cfg`` with the namelist defaults of the reference
    (code/mpi-io/OCFD_readpara.f90:65-110) for fields not given."""
    c = OcfdCfg()
    c.nx, c.ny, c.nz, c.lap, c.n_spec = nx, ny, nz, LAP, n_spec
    for a in range(3):
        c.periodic[a] = int(periodic[a])
    c.scheme_invis, c.scheme_vis, c.flux_splitting = scheme_invis, scheme_vis, flux_splitting
    c.if_viscous, c.if_scheme_character, c.ibc = if_viscous, if_scheme_character, ibc
    for a in range(6):
        c.scheme_boundary[a] = 0 if scheme_boundary is None else int(scheme_boundary[a])
        c.ghost_cell[a] = 0
    c.if_mass_force = 1 if mass_force is not None else 0
    c.if_real_gas = if_real_gas
    c.flag_vis, c.flag_cond, c.flag_diff = flag_vis, flag_cond, flag_diff
    c.wall_catalysis = 0
    c.ud7l_diss = ud7l_diss
    hp = [1.0] * 16 if hybrid_para is None else list(hybrid_para) + [1.0] * (16 - len(hybrid_para))
    for a in range(16):
        c.hybrid_para[a] = hp[a]
    c.amu0, c.pr, c.sc = amu0, pr, sc
    mf = mass_force or (0.0, 0.0, 0.0)
    for a in range(3):
        c.mass_force[a] = mf[a]
    bp = list(bc_para or []) + [0.0] * (32 - len(bc_para or []))
    for a in range(32):
        c.bc_para[a] = bp[a]
    return c


_LIB = None


def load_library(path=None):
    """Load the HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    i32, dbl, vp = ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
    P = ctypes.POINTER
    sig = {
        "ocfd_init": (i32, [P(OcfdCfg)]), "ocfd_finalize": (i32, []),
        "ocfd_last_error": (ctypes.c_char_p, []), "ocfd_set_option": (i32, [i32, i32]),
        "ocfd_set_species": (i32, [i32, vp, dbl, vp, vp]),
        "ocfd_set_mesh": (i32, [vp, vp, vp, vp]),
        "ocfd_upload_field": (i32, [i32, vp]), "ocfd_download_field": (i32, [i32, vp]),
        "ocfd_download_mask": (i32, [i32, vp]),
        "ocfd_set_bc_boundarylayer": (i32, [vp, vp, vp, vp]), "ocfd_set_bc_time": (i32, [dbl]),
        "ocfd_flow_time_advance": (i32, [dbl]), "ocfd_save_stage0": (i32, []),
        "ocfd_stage_begin": (i32, [i32]), "ocfd_comput_amu_amk": (i32, []),
        "ocfd_comput_rhybrid": (i32, []), "ocfd_du_inviscous": (i32, []), "ocfd_du_viscous": (i32, []),
        "ocfd_time_adv_rk3": (i32, [i32, dbl]), "ocfd_comput_flow_variables": (i32, []),
        "ocfd_bc": (i32, []), "ocfd_negative_et_count": (i32, []),
        "ocfd_filtering": (i32, [i32, i32, vp]),
        "ocfd_hybrid_certificate": (i32, [P(ctypes.c_int64), P(dbl), P(dbl)]),
        "ocfd_metric_tables_in_use": (i32, []),
        "ocfd_show_flow_msg": (i32, [vp]), "ocfd_ana_box": (i32, [vp]),
        "ocfd_read_flow_data": (i32, [ctypes.c_char_p, P(i32), P(dbl)]),
        "ocfd_save_flow_data": (i32, [ctypes.c_char_p, i32, dbl]),
        "ocfd_read_mesh": (i32, [ctypes.c_char_p, i32, vp]),
        "ocfd_flux1d_batch": (i32, [i32, i32, ctypes.c_int64, i32, vp, vp, vp, P(i32), P(i32)]), "ocfd_synchronize": (i32, []),
        "ocfd_stream": (vp, []), "ocfd_kernel_count": (i32, []), "ocfd_kernel_name": (ctypes.c_char_p, [i32]),
        "ocfd_kernel_stats": (i32, [i32, P(dbl), P(ctypes.c_int64)]), "ocfd_kernel_stats_reset": (i32, []),
        "ocfd_autotune": (i32, [i32]), "ocfd_set_tile": (i32, [i32, i32]),
        "ocfd_tile_choice": (i32, [i32, P(i32), P(i32), P(ctypes.c_float)]),
        "ocfd_set_transport": (i32, [vp, vp, vp]), "ocfd_set_bc_jet": (i32, [vp]),
        "ocfd_ns_step": (i32, [dbl, i32, i32, i32]), "ocfd_chemical_time_advance": (i32, [dbl, i32]),
        "ocfd_comput_T_P": (i32, []), "ocfd_du_species": (i32, []), "ocfd_species_time_adv": (i32, [i32, dbl]),
        "ocfd_du_scaler_invis": (i32, [i32]), "ocfd_du_scaler_viscous": (i32, [i32]),
        "ocfd_scalar_time_adv_rk3": (i32, [i32, i32, dbl]),
        "ocfd_set_chem_method": (i32, [i32, dbl, dbl, dbl, i32, i32]), "ocfd_cfl_dt": (i32, [dbl, P(dbl)]),
        "ocfd_chem_set_mechanism": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, dbl, vp, vp]),
        "ocfd_chem_set_reverse_rates": (i32, [vp]),
        "ocfd_chem_timeadv_batch": (i32, [ctypes.c_int64, vp, vp, vp, i32, i32]),
        "ocfd_chem_adaptive_batch": (i32, [ctypes.c_int64, vp, vp, vp, dbl, dbl, dbl, i32, vp, i32]),
        "ocfd_chem_adaptive_batch_noalloc": (i32, [ctypes.c_int64, vp, vp, vp, dbl, dbl, dbl, i32, vp]),
        "ocfd_chem_synchronize": (i32, []), "ocfd_chem_stream": (vp, []), "ocfd_chem_set_stream": (i32, [vp]),
        "ocfd_chem_last_error": (ctypes.c_char_p, []), "ocfd_chem_unfinished": (ctypes.c_int64, []),
        "ocfd_chem_counters": (i32, [vp, i32]), "ocfd_chem_set_option": (i32, [i32, i32]),
        "ocfd_mlp_create": (i32, [i32, vp]), "ocfd_mlp_destroy": (i32, []), "ocfd_mlp_param_count": (ctypes.c_int64, []),
        "ocfd_mlp_set_params": (i32, [vp]), "ocfd_mlp_get_params": (i32, [vp]), "ocfd_mlp_set_stream": (i32, [vp]),
        "ocfd_mlp_forward": (i32, [ctypes.c_int64, vp, vp]),
        "ocfd_mlp_gradients": (i32, [ctypes.c_int64, vp, vp, vp, vp]),
        "ocfd_mlp_train_step": (i32, [ctypes.c_int64, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, vp]),
        "ocfd_mlp_synchronize": (i32, []), "ocfd_mlp_last_error": (ctypes.c_char_p, []),
        "ocfd_mlp_gemm_bf16": (i32, [i32, i32, i32, i32, i32, vp, ctypes.c_int64, vp, ctypes.c_int64, vp,
                                     ctypes.c_int64, i32]),
        "ocfd_chem_surrogate_set": (i32, [vp, i32]),
        "ocfd_chem_surrogate_batch": (i32, [ctypes.c_int64, vp, vp, vp, i32, dbl, dbl, dbl, i32, vp]),
        "ocfd_chem_surrogate_fallbacks": (ctypes.c_int64, []),
        "ocfd_chem_surrogate_autotune": (i32, [ctypes.c_int64, vp, vp, vp, i32, vp]),
        "ocfd_chem_surrogate_tile": (i32, [i32, P(i32), P(i32)]),
        "ocfd_chem_cell_features": (i32, [ctypes.c_int64, vp, vp, vp, i32, vp]),
        "ocfd_chem_surrogate_targets": (i32, [ctypes.c_int64, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    if path is None:
        _LIB = lib
    return lib


def filter_array(filters):
    """Para%Filter(100,10) (column-major) holding the given rows of
    Filter(1:13,m); unused entries are zero."""
    if len(filters) > 10:
        raise ValueError("at most 10 filters (NFiltering <= 10)")
    F = np.zeros((100, 10), dtype=np.float64, order="F")
    for m, row in enumerate(filters):
        row = np.asarray(row, dtype=np.float64)
        if row.shape != (13,):
            raise ValueError("a filter row holds the 13 values Filter(1:13,m)")
        F[:13, m] = row
    return F


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OcfdError(RuntimeError):
    pass


def bl_data_shapes(cfg):
    """Shapes of the boundary-layer inlet and upper data (Fortran order,
    variable index last) for BC_Para(1) (BcData_inlet 0, 1, 2) and
    BC_Para(2) (BcData_upper -1, 0, 1, 2): inlet (nv,), (ny, nv) or
    (ny, nz, nv); upper None, (nv,), (nx, nv) or (nx, nz, nv), nv = 5 + N_SPEC."""
    nv = 5 + cfg.n_spec
    ki, ku = int(round(cfg.bc_para[0])), int(round(cfg.bc_para[1]))
    inl = {0: (nv,), 1: (cfg.ny, nv), 2: (cfg.ny, cfg.nz, nv)}.get(ki)
    up = {-1: None, 0: (nv,), 1: (cfg.nx, nv), 2: (cfg.nx, cfg.nz, nv)}.get(ku, "bad")
    if inl is None or up == "bad":
        raise ValueError(f"BC_Para(1) = {ki} / BC_Para(2) = {ku}: inlet data 0, 1, 2 and upper data -1, 0, 1, 2")
    return inl, up


def bl_data_arrays(cfg, inlet, upper):
    """inlet and upper as Fortran-ordered float64 arrays of the shapes
    ``bl_data_shapes`` gives (upper is None for BC_Para(2) = -1)."""
    s_in, s_up = bl_data_shapes(cfg)
    inl = np.asfortranarray(inlet, dtype=np.float64)
    if inl.shape != s_in:
        raise ValueError(f"inlet data: expected shape {s_in}, got {inl.shape}")
    if s_up is None:
        return inl, None
    up = np.asfortranarray(upper, dtype=np.float64)
    if up.shape != s_up:
        raise ValueError(f"upper data: expected shape {s_up}, got {up.shape}")
    return inl, up


class FlowSolver:
    """Device-resident flow state and the hot-path entry points.

    Method names follow the reference subroutines; every call goes through
    the HIP library (there is no host fallback)."""

    def __init__(self, cfg, lib=None):
        self.lib = lib or load_library()
        self.cfg = cfg
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        self.halo_shape = (cfg.nx + 2 * LAP, cfg.ny + 2 * LAP, cfg.nz + 2 * LAP)
        self._check(self.lib.ocfd_init(ctypes.byref(cfg)))

    # -- plumbing -------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise OcfdError(self.lib.ocfd_last_error().decode())

    def close(self):
        self.lib.ocfd_finalize()

    def set_option(self, key, value):
        self._check(self.lib.ocfd_set_option(key, int(value)))

    def field_shape(self, name):
        if name in ("F", "FN", "DU"):
            return (self.nx, self.ny, self.nz, 5)
        if name in ("AMU", "AMK"):
            return (self.nx, self.ny, self.nz)
        if name in ("DIN", "AMD", "DUS"):
            return (self.nx, self.ny, self.nz, self.cfg.n_spec)
        if name == "QP":
            return (self.nx, self.ny, self.nz, 3)
        if name == "DI":
            return self.halo_shape + (self.cfg.n_spec,)
        return self.halo_shape

    def upload(self, name, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != self.field_shape(name):
            raise ValueError(f"{name}: expected shape {self.field_shape(name)}, got {a.shape}")
        self._check(self.lib.ocfd_upload_field(FIELD[name], _ptr(a)))

    def download(self, name):
        a = np.zeros(self.field_shape(name), dtype=np.float64, order="F")
        self._check(self.lib.ocfd_download_field(FIELD[name], _ptr(a)))
        return a

    def download_mask(self, direction):
        shp = [self.nx, self.ny, self.nz]
        shp[direction] += 1
        a = np.zeros(tuple(shp), dtype=np.uint8, order="F")
        self._check(self.lib.ocfd_download_mask(direction, _ptr(a)))
        return a

    # -- setup ----------------------------------------------------------
    def set_species(self, mi, n_tct=1, tct=None, tct_max=5000.0, coef=None):
        mi = np.ascontiguousarray(mi, dtype=np.float64)
        t = None if tct is None else np.ascontiguousarray(tct, dtype=np.float64)
        cf = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        self._check(self.lib.ocfd_set_species(n_tct, None if t is None else _ptr(t), tct_max, _ptr(mi),
                                              None if cf is None else _ptr(cf)))

    def set_mesh(self, axx, ayy, azz, span=None):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (axx, ayy, azz)]
        sp = np.zeros(9) if span is None else np.ascontiguousarray(span, dtype=np.float64)
        self._check(self.lib.ocfd_set_mesh(_ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(sp)))

    def set_bc_boundarylayer(self, inlet, upper, nonreflect_upper, wall_pertb):
        """BC%flow_inlet0d/1d/2d and BC%flow_upper0d/1d/2d
        (init_bcdata_boundary, OCFD_BC_Boundarylayers.f90:247-430) by
        BC_Para(1) / BC_Para(2); shapes in ``bl_data_shapes``."""
        inl, up = bl_data_arrays(self.cfg, inlet, upper)
        nr = np.asfortranarray(nonreflect_upper, dtype=np.int32)
        wp = np.asfortranarray(wall_pertb, dtype=np.float64)
        if nr.shape != (self.nx, self.nz) or wp.shape != (self.nx, self.nz):
            raise ValueError("nonreflect_upper and wall_pertb must have shape (nx, nz)")
        self._check(self.lib.ocfd_set_bc_boundarylayer(_ptr(inl), None if up is None else _ptr(up), _ptr(nr),
                                                       _ptr(wp)))

    def set_bc_time(self, ht):
        self._check(self.lib.ocfd_set_bc_time(float(ht)))

    def set_transport(self, mu_an=None, k_an=None, D_an=None):
        """Chemkin fits (n_spec, 4), (n_spec, 4), (n_spec, n_spec, 4) (pair k <= j)."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (mu_an, k_an, D_an)]
        self._check(self.lib.ocfd_set_transport(*[None if a is None else _ptr(a) for a in arrs]))

    def set_bc_jet(self, inlet2d):
        """flow_inlet2d of shape (ny, nz, 5 + n_spec): d, u, v, w, T, di."""
        a = np.asfortranarray(inlet2d, dtype=np.float64)
        if a.shape != (self.ny, self.nz, 5 + self.cfg.n_spec):
            raise ValueError(f"inlet2d: expected shape {(self.ny, self.nz, 5 + self.cfg.n_spec)}, got {a.shape}")
        self._check(self.lib.ocfd_set_bc_jet(_ptr(a)))

    def read_flow_data(self, path):
        """read_flow_data + init3d from a restart file of the reference's
        layout; returns (Istep, tt).  Follow with comput_flow_variables and
        OCFD_bc, as NS_solver does."""
        istep, tt = ctypes.c_int32(0), ctypes.c_double(0.0)
        self._check(self.lib.ocfd_read_flow_data(os.fsencode(path), ctypes.byref(istep), ctypes.byref(tt)))
        return int(istep.value), float(tt.value)

    def save_flow_data(self, path, istep, tt):
        """save_flow_data: d, u, v, w, T, di of the resident state."""
        self._check(self.lib.ocfd_save_flow_data(os.fsencode(path), int(istep), float(tt)))

    def read_mesh(self, path, gridtype, span=None):
        """init_mesh from an OCFD-grid.dat of Iflag_Gridtype 10, 20, 21, 30 or 31."""
        sp = np.zeros(9) if span is None else np.ascontiguousarray(span, dtype=np.float64)
        self._check(self.lib.ocfd_read_mesh(os.fsencode(path), int(gridtype), _ptr(sp)))

    def upload_state(self, state):
        """state: dict with F, DI, D, U, V, W, T, P, ET (reference layouts)."""
        for name in ("F", "DI", "D", "U", "V", "W", "T", "P", "ET"):
            self.upload(name, state[name])

    # -- hot path (reference names) -------------------------------------
    def flow_time_advance(self, dt1):
        self._check(self.lib.ocfd_flow_time_advance(float(dt1)))

    def NS_step(self, dt, nsub=1, mode=1, reaction=True):
        """One time step of NS_solver (OCFD_NS_Solver.f90:26-37)."""
        self._check(self.lib.ocfd_ns_step(float(dt), int(nsub), int(mode), int(bool(reaction))))

    def cfl_dt(self, cfl):
        """dt = cfl / max sum_dir (|u_n| + c |grad xi|) / h (device reduction)."""
        out = ctypes.c_double(0.0)
        self._check(self.lib.ocfd_cfl_dt(float(cfl), ctypes.byref(out)))
        return out.value

    def set_chem_method(self, method, rtol_y=0.05, y_floor=1e-6, rtol_t=0.01, max_sub=4000, scheduler=1):
        """CHEM_FIXED (chemical_timeAdv, CHEM_TimeAdv = mode), CHEM_ADAPTIVE or
        CHEM_SURROGATE (with the adaptive integrator as fallback)."""
        self._check(self.lib.ocfd_set_chem_method(int(method), float(rtol_y), float(y_floor), float(rtol_t),
                                                  int(max_sub), int(scheduler)))

    def Chemical_time_advance(self, dt2, mode=1):
        self._check(self.lib.ocfd_chemical_time_advance(float(dt2), int(mode)))

    def comput_T_P(self):
        self._check(self.lib.ocfd_comput_T_P())

    def du_species(self):
        """du_scaler_invis + du_scaler_viscous of every species (into DUS, Qp into QP)."""
        self._check(self.lib.ocfd_du_species())

    def Scalar_time_adv_RK3(self, krk, dt):
        """Species right-hand sides and update of stage krk (OCFD_Time_Adv.f90:64-71)."""
        self._check(self.lib.ocfd_species_time_adv(int(krk), float(dt)))

    def du_scaler_invis(self, m):
        """Inviscid right-hand side of species m (1-based) into DUS[..., m-1]
        (du_scaler_invis, OCFD_scaler_inviscous.f90:1-84)."""
        self._check(self.lib.ocfd_du_scaler_invis(int(m)))

    def du_scaler_viscous(self, m):
        """DUS[..., m-1] += diffusive right-hand side of species m; Qp zeroed at
        m = 1 (du_scaler_viscous, OCFD_scaler_viscous.f90:4-79)."""
        self._check(self.lib.ocfd_du_scaler_viscous(int(m)))

    def scalar_time_adv_rk3(self, krk, m, dt):
        """Scalar_time_adv_RK3(KRK, m, dus, dt) of species m (OCFD_Time_Adv.f90:134-166)."""
        self._check(self.lib.ocfd_scalar_time_adv_rk3(int(krk), int(m), float(dt)))

    def save_stage0(self):
        self._check(self.lib.ocfd_save_stage0())

    def stage_begin(self, krk):
        self._check(self.lib.ocfd_stage_begin(int(krk)))

    def comput_Amu_Amk_AMD(self):
        self._check(self.lib.ocfd_comput_amu_amk())

    def comput_Rhybrid(self):
        self._check(self.lib.ocfd_comput_rhybrid())

    def du_inviscous(self):
        self._check(self.lib.ocfd_du_inviscous())

    def du_viscous(self):
        self._check(self.lib.ocfd_du_viscous())

    def OCFD_time_adv_RK3(self, krk, dt):
        self._check(self.lib.ocfd_time_adv_rk3(int(krk), float(dt)))

    def comput_flow_variables(self):
        self._check(self.lib.ocfd_comput_flow_variables())

    def OCFD_bc(self):
        self._check(self.lib.ocfd_bc())

    def filtering(self, istep, filters):
        """filtering (OCFD_filtering.f90:4-76) after Istep = istep.  filters:
        a sequence of up to 10 rows of the 13 Filter(1:13,m) values (step
        interval, X/Y/Z flags, ib, ie, jb, je, kb, ke, scheme 1 = 9-point or
        2 = shock-capturing, s0, rth), passed as the reference's
        Para%Filter(100,10)."""
        F = filter_array(filters)
        self._check(self.lib.ocfd_filtering(int(istep), len(filters), _ptr(F)))

    def OCFD2d_flux(self, side, scheme, v, bound_index=(0, 0), scheme_boundary=(0, 0), scm=None):
        """OCFD2d_flux1 (side 0) / OCFD2d_flux2 (side 1) over a batch of
        pencils (OCFD_inviscous.f90:209-362).  v is a float64 CUDA tensor of
        shape (npen, n + 2*LAP) holding f1d(1-LAP:n+LAP); scm, for the hybrid
        scheme, a uint8 CUDA tensor of shape (npen, n + 1).  Returns hh(0:n)
        per pencil as a (npen, n + 1) float64 CUDA tensor."""
        import torch
        if not (isinstance(v, torch.Tensor) and v.is_cuda and v.dtype == torch.float64 and v.dim() == 2):
            raise TypeError("v must be a 2-D float64 CUDA tensor")
        v = v.contiguous()
        npen, n = v.shape[0], v.shape[1] - 2 * LAP
        if n < 1:
            raise ValueError("v needs more than 2*LAP points per pencil")
        hh = torch.empty((npen, n + 1), dtype=torch.float64, device=v.device)
        sp = 0
        if scm is not None:
            if not (isinstance(scm, torch.Tensor) and scm.is_cuda and scm.dtype == torch.uint8):
                raise TypeError("scm must be a uint8 CUDA tensor")
            if tuple(scm.shape) != (npen, n + 1):
                raise ValueError(f"scm must have shape {(npen, n + 1)}")
            scm = scm.contiguous()
            sp = scm.data_ptr()
        b = (ctypes.c_int32 * 2)(*bound_index)
        sb = (ctypes.c_int32 * 2)(*scheme_boundary)
        # the library runs on its own stream: order it after torch's work
        torch.cuda.current_stream().synchronize()
        self._check(self.lib.ocfd_flux1d_batch(int(side), int(scheme), npen, n, v.data_ptr(), hh.data_ptr(), sp, b, sb))
        self._check(self.lib.ocfd_synchronize())
        return hh

    ANA_BOX_NAMES = ("Re_lamda", "Re_lamdax", "Amt", "vv_rms", "u_rms", "v_rms", "w_rms", "u_skewness",
                     "u_flatness", "ux_skewness", "ux_flatness", "E_kinetic", "d_rms", "T_aver", "Amu_aver",
                     "d_aver", "Ux_rms", "Alamdax")

    def show_flow_msg(self):
        """(mean total energy f5/J, mean kinetic energy) over the interior
        (show_flow_msg, OCFD_Time_Adv.f90:200-239)."""
        e = np.zeros(2)
        self._check(self.lib.ocfd_show_flow_msg(_ptr(e)))
        return float(e[0]), float(e[1])

    def ana_box(self):
        """Isotropic-turbulence statistics of ana_box (OCFD_ANA.f90:197-303)
        as a dict keyed by the reference's statistics.dat column names."""
        out = np.zeros(18)
        self._check(self.lib.ocfd_ana_box(_ptr(out)))
        return dict(zip(self.ANA_BOX_NAMES, (float(x) for x in out)))

    def hybrid_certificate(self):
        """(uncertain interfaces, smallest |R/theta - 1|, band eta) of the
        last certified fast sensor mean; uncertain == 0 certifies that the
        fast masks are the ordered ones (otherwise the library recomputed
        them from the ordered mean)."""
        n, margin, eta = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.ocfd_hybrid_certificate(ctypes.byref(n), ctypes.byref(margin), ctypes.byref(eta)))
        return int(n.value), float(margin.value), float(eta.value)

    def negative_et_count(self):
        return int(self.lib.ocfd_negative_et_count())

    def metric_tables_in_use(self):
        """True when the stepping kernels evaluate the metrics from the
        per-row coordinate-derivative tables (OPT_METRIC_TABLES)."""
        return bool(self.lib.ocfd_metric_tables_in_use())

    def synchronize(self):
        self._check(self.lib.ocfd_synchronize())

    # -- tile autotuner -------------------------------------------------
    def autotune(self, reps=3):
        """Time every sweep tile shape on the current state and keep the
        fastest per direction; returns {dir: (variant, NP, TS, [ms...])}."""
        self._check(self.lib.ocfd_autotune(int(reps)))
        return {d: self.tile_choice(d) for d in range(3)}

    def tile_choice(self, direction):
        np_, ts = ctypes.c_int32(), ctypes.c_int32()
        ms = (ctypes.c_float * 4)()
        v = self.lib.ocfd_tile_choice(direction, ctypes.byref(np_), ctypes.byref(ts), ms)
        return v, np_.value, ts.value, [float(x) for x in ms]

    def set_tile(self, direction, variant):
        self._check(self.lib.ocfd_set_tile(direction, variant))

    # -- timing ---------------------------------------------------------
    def kernel_stats(self):
        out = {}
        for i in range(self.lib.ocfd_kernel_count()):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            self._check(self.lib.ocfd_kernel_stats(i, ctypes.byref(ms), ctypes.byref(n)))
            out[self.lib.ocfd_kernel_name(i).decode()] = (ms.value, n.value)
        return out

    def kernel_stats_reset(self):
        self._check(self.lib.ocfd_kernel_stats_reset())


__all__ = ["OcfdCfg", "make_cfg", "load_library", "FlowSolver", "OcfdError", "FIELD", "FIELDS", "LAP"]
