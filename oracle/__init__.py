"""CPU oracle for the flow and chemistry paths (TEST INFRASTRUCTURE ONLY).

A ctypes wrapper around liboracle.so (built from ocfd_oracle*.c by the
Makefile in this directory).  ``Oracle`` mirrors the method names of the
product's ``FlowSolver`` so a test can run the same sequence on both.  Only
tests/, the smoke check and the CPU-baseline leg of bench.py import this.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_LIB = None


def load(ocfd_module):
    """Load liboracle.so; ``ocfd_module`` is the product package (for the
    shared ocfd_cfg layout)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    i32, dbl, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    P = ctypes.POINTER
    sig = {
        "orc_init": (i32, [P(ocfd_module.OcfdCfg)]), "orc_finalize": (i32, []),
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_set_species": (i32, [i32, vp, dbl, vp, vp]), "orc_set_mesh": (i32, [vp, vp, vp, vp]),
        "orc_upload_field": (i32, [i32, vp]), "orc_download_field": (i32, [i32, vp]),
        "orc_download_mask": (i32, [i32, vp]),
        "orc_set_bc_boundarylayer": (i32, [vp, vp, vp, vp]), "orc_set_bc_time": (i32, [dbl]),
        "orc_flow_time_advance": (i32, [dbl]), "orc_comput_amu_amk": (i32, []),
        "orc_du_inviscous": (i32, []), "orc_du_viscous": (i32, []), "orc_time_adv_rk3": (i32, [i32, dbl]),
        "orc_comput_flow_variables": (i32, []), "orc_bc": (i32, []), "orc_comput_rhybrid": (i32, []),
        "orc_exchange_all": (i32, []), "orc_negative_et_count": (i32, []), "orc_sensor_mean": (dbl, []),
        "orc_show_flow_msg": (i32, [vp]), "orc_ana_box": (i32, [vp]),
        "orc_dx0": (None, [vp, vp]), "orc_dy0": (None, [vp, vp]), "orc_dz0": (None, [vp, vp]),
        "orc_recon_flux1": (i32, [i32, vp, vp, i32, vp, vp, vp]),
        "orc_recon_flux2": (i32, [i32, vp, vp, i32, vp, vp, vp]),
        "orc_set_ud7l_diss": (i32, [dbl]),
        "orc_split": (i32, [i32, i32] + [vp] * 12),
        "orc_set_transport": (i32, [vp, vp, vp]), "orc_set_bc_jet": (i32, [vp]),
        "orc_ns_step": (i32, [dbl, i32, i32, i32]), "orc_chemical_time_advance": (i32, [dbl, i32]),
        "orc_comput_T_P": (i32, []), "orc_du_species": (i32, [vp]), "orc_download_qp": (i32, [vp]),
        "orc_char_matrices": (i32, [vp, vp, vp]),
        "orc_comput_T": (dbl, [vp, dbl]), "orc_comput_E": (dbl, [vp, dbl]), "orc_comput_Cp": (dbl, [dbl, vp, dbl]),
        "orc_set_threads": (i32, [i32]),
        "orc_filtering": (i32, [i32, i32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _LIB = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """Same interface as the product's FlowSolver, computed by the C restatement."""

    def __init__(self, ocfd_module, cfg):
        self.m = ocfd_module
        self.lib = load(ocfd_module)
        self.cfg = cfg
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        L = ocfd_module.LAP
        self.halo_shape = (cfg.nx + 2 * L, cfg.ny + 2 * L, cfg.nz + 2 * L)
        self._check(self.lib.orc_init(ctypes.byref(cfg)))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def filtering(self, istep, filters):
        """filtering (OCFD_filtering.f90:4-76) after Istep = istep; filters as
        the product's FlowSolver.filtering."""
        F = self.m.filter_array(filters)
        self._check(self.lib.orc_filtering(int(istep), len(filters), _ptr(F)))

    def set_threads(self, n):
        """OpenMP threads of the flow loops (process-wide); returns the count
        in use.  Results do not depend on it: the loops that run in parallel
        are per cell or per pencil, and the sums stay sequential."""
        return self.lib.orc_set_threads(int(n))

    field_shape = None  # bound below

    def upload(self, name, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        assert a.shape == self.field_shape(name), (name, a.shape)
        self._check(self.lib.orc_upload_field(self.m.FIELD[name], _ptr(a)))

    def download(self, name):
        a = np.zeros(self.field_shape(name), dtype=np.float64, order="F")
        self._check(self.lib.orc_download_field(self.m.FIELD[name], _ptr(a)))
        return a

    def download_mask(self, direction):
        shp = [self.nx, self.ny, self.nz]
        shp[direction] += 1
        a = np.zeros(tuple(shp), dtype=np.uint8, order="F")
        self._check(self.lib.orc_download_mask(direction, _ptr(a)))
        return a

    def set_species(self, mi, n_tct=1, tct=None, tct_max=5000.0, coef=None):
        mi = np.ascontiguousarray(mi, dtype=np.float64)
        t = None if tct is None else np.ascontiguousarray(tct, dtype=np.float64)
        cf = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        self._check(self.lib.orc_set_species(n_tct, None if t is None else _ptr(t), tct_max, _ptr(mi),
                                             None if cf is None else _ptr(cf)))

    def set_mesh(self, axx, ayy, azz, span=None):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (axx, ayy, azz)]
        sp = np.zeros(9) if span is None else np.ascontiguousarray(span, dtype=np.float64)
        self._check(self.lib.orc_set_mesh(_ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(sp)))

    def set_bc_boundarylayer(self, inlet, upper, nonreflect_upper, wall_pertb):
        inl, up = self.m.bl_data_arrays(self.cfg, inlet, upper)
        nr = np.asfortranarray(nonreflect_upper, dtype=np.int32)
        wp = np.asfortranarray(wall_pertb, dtype=np.float64)
        if nr.shape != (self.nx, self.nz) or wp.shape != (self.nx, self.nz):
            raise ValueError("nonreflect_upper and wall_pertb must have shape (nx, nz)")
        self._check(self.lib.orc_set_bc_boundarylayer(_ptr(inl), None if up is None else _ptr(up), _ptr(nr),
                                                      _ptr(wp)))

    def set_bc_time(self, ht):
        self._check(self.lib.orc_set_bc_time(float(ht)))

    def upload_state(self, state):
        for name in ("F", "DI", "D", "U", "V", "W", "T", "P", "ET"):
            self.upload(name, state[name])

    def flow_time_advance(self, dt1):
        self._check(self.lib.orc_flow_time_advance(float(dt1)))

    def comput_flow_variables(self):
        self._check(self.lib.orc_comput_flow_variables())

    def OCFD_bc(self):
        self._check(self.lib.orc_bc())

    def comput_Amu_Amk_AMD(self):
        self._check(self.lib.orc_comput_amu_amk())

    def du_inviscous(self):
        self._check(self.lib.orc_du_inviscous())

    def du_viscous(self):
        self._check(self.lib.orc_du_viscous())

    def OCFD_time_adv_RK3(self, krk, dt):
        self._check(self.lib.orc_time_adv_rk3(int(krk), float(dt)))

    def comput_Rhybrid(self):
        self._check(self.lib.orc_comput_rhybrid())

    def exchange_all(self):
        self._check(self.lib.orc_exchange_all())

    def negative_et_count(self):
        return int(self.lib.orc_negative_et_count())

    def show_flow_msg(self):
        e = np.zeros(2)
        self._check(self.lib.orc_show_flow_msg(_ptr(e)))
        return float(e[0]), float(e[1])

    def ana_box(self):
        out = np.zeros(18)
        self._check(self.lib.orc_ana_box(_ptr(out)))
        return dict(zip(self.m.FlowSolver.ANA_BOX_NAMES, (float(x) for x in out)))

    def sensor_mean(self):
        return float(self.lib.orc_sensor_mean())

    def set_transport(self, mu_an=None, k_an=None, D_an=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (mu_an, k_an, D_an)]
        self._check(self.lib.orc_set_transport(*[None if a is None else _ptr(a) for a in arrs]))

    def set_bc_jet(self, inlet2d):
        a = np.asfortranarray(inlet2d, dtype=np.float64)
        assert a.shape == (self.ny, self.nz, 5 + self.cfg.n_spec)
        self._check(self.lib.orc_set_bc_jet(_ptr(a)))

    def NS_step(self, dt, nsub=1, mode=1, reaction=True):
        self._check(self.lib.orc_ns_step(float(dt), int(nsub), int(mode), int(bool(reaction))))

    def Chemical_time_advance(self, dt2, mode=1):
        self._check(self.lib.orc_chemical_time_advance(float(dt2), int(mode)))

    def comput_T_P(self):
        self._check(self.lib.orc_comput_T_P())

    def du_species(self):
        """(DUS (nx, ny, nz, n_spec), QP (nx, ny, nz, 3)) of the current state."""
        dus = np.zeros((self.nx, self.ny, self.nz, self.cfg.n_spec), order="F")
        qp = np.zeros((self.nx, self.ny, self.nz, 3), order="F")
        self._check(self.lib.orc_du_species(_ptr(dus)))
        self._check(self.lib.orc_download_qp(_ptr(qp)))
        return dus, qp

    def close(self):
        self.lib.orc_finalize()


def _field_shape(self, name):
    if name in ("F", "FN", "DU"):
        return (self.nx, self.ny, self.nz, 5)
    if name in ("AMU", "AMK"):
        return (self.nx, self.ny, self.nz)
    if name in ("DIN", "AMD"):
        return (self.nx, self.ny, self.nz, self.cfg.n_spec)
    if name == "DI":
        return self.halo_shape + (self.cfg.n_spec,)
    return self.halo_shape


Oracle.field_shape = _field_shape


class ChemOracle:
    """Chemistry restatement (CHEM_reactions.f90 / CHEM_thermdyn.f90)."""

    def __init__(self, ocfd_module, mech):
        self.lib = load(ocfd_module)
        L = self.lib
        i32, dbl, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        sig = {
            "orc_chem_set_mechanism": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, dbl, vp, vp]),
            "orc_chem_E": (dbl, [vp, dbl]), "orc_chem_T": (dbl, [vp, dbl]), "orc_chem_gibbs": (dbl, [i32, dbl]),
            "orc_chem_rate_coef": (i32, [vp, dbl, vp, vp, vp]), "orc_chem_rates": (i32, [vp, dbl, vp, vp]),
            "orc_gauss": (i32, [i32, vp, vp, vp]), "orc_chem_timeadv": (i32, [i32, vp, vp, dbl]),
            "orc_chem_batch": (i32, [i32, ctypes.c_long, vp, vp, vp, i32]),
            "orc_chem_adaptive": (i32, [vp, vp, dbl, dbl, dbl, dbl, i32, vp, vp]),
            "orc_chem_adaptive_batch": (i32, [ctypes.c_long, vp, vp, vp, dbl, dbl, dbl, i32, vp, vp]),
            "orc_set_threads": (i32, [i32]), "orc_chem_set_reverse": (i32, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        self.m = mech
        self._keep = [np.ascontiguousarray(a) for a in (
            mech.af_type.astype(np.int32), mech.third_body.astype(np.int32), mech.fc_troe.astype(np.float64),
            mech.arrhenius.astype(np.float64), mech.vf.astype(np.int32), mech.vr.astype(np.int32),
            mech.eff.astype(np.float64), np.asarray(mech.tct, dtype=np.float64), mech.mi.astype(np.float64),
            mech.coef.astype(np.float64))]
        k = self._keep
        rc = L.orc_chem_set_mechanism(mech.n_spec, mech.n_reac, mech.rereac, _ptr(k[0]), _ptr(k[1]), _ptr(k[2]),
                                      _ptr(k[3]), _ptr(k[4]), _ptr(k[5]), _ptr(k[6]), mech.n_tct, _ptr(k[7]),
                                      float(mech.tct_max), _ptr(k[8]), _ptr(k[9]))
        if rc:
            raise RuntimeError("orc_chem_set_mechanism failed")
        if mech.rereac != 0:
            self._rev = np.ascontiguousarray(mech.reverse, dtype=np.float64)
            L.orc_chem_set_reverse(_ptr(self._rev))

    def set_threads(self, n):
        """OpenMP threads of the batch functions; returns the count in use."""
        return self.lib.orc_set_threads(int(n))

    def E(self, di, T):
        return self.lib.orc_chem_E(_ptr(np.ascontiguousarray(di, dtype=np.float64)), float(T))

    def T(self, di, E):
        return self.lib.orc_chem_T(_ptr(np.ascontiguousarray(di, dtype=np.float64)), float(E))

    def gibbs(self, k, T):
        return self.lib.orc_chem_gibbs(int(k), float(T))

    def rate_coef(self, ci, T):
        nr = self.m.n_reac
        wf, wr, kj = np.zeros(nr), np.zeros(nr), np.zeros(nr)
        c = np.ascontiguousarray(ci, dtype=np.float64)
        self.lib.orc_chem_rate_coef(_ptr(c), float(T), _ptr(wf), _ptr(wr), _ptr(kj))
        return wf, wr, kj

    def rates(self, ci, T, jacobian=True):
        ns = self.m.n_spec
        ri = np.zeros(ns)
        A = np.zeros((ns, ns)) if jacobian else None
        c = np.ascontiguousarray(ci, dtype=np.float64)
        self.lib.orc_chem_rates(_ptr(c), float(T), _ptr(ri), _ptr(A) if jacobian else None)
        return ri, A

    def gauss(self, A, b):
        A = np.ascontiguousarray(A, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        x = np.zeros(len(b))
        self.lib.orc_gauss(len(b), _ptr(A), _ptr(b), _ptr(x))
        return x

    def timeadv(self, mode, di, E, dt):
        d = np.ascontiguousarray(di, dtype=np.float64).copy()
        e = np.array([E], dtype=np.float64)
        if self.lib.orc_chem_timeadv(int(mode), _ptr(d), _ptr(e), float(dt)):
            raise RuntimeError("orc_chem_timeadv failed")
        return d, float(e[0])

    def batch(self, mode, rho, E, dt, nsub):
        """rho [n][ns] (C order), E [n], dt [n]; returns new copies."""
        r = np.ascontiguousarray(rho, dtype=np.float64).copy()
        e = np.ascontiguousarray(E, dtype=np.float64).copy()
        d = np.ascontiguousarray(dt, dtype=np.float64)
        if self.lib.orc_chem_batch(int(mode), len(e), _ptr(r), _ptr(e), _ptr(d), int(nsub)):
            raise RuntimeError("orc_chem_batch failed")
        return r, e

    def adaptive_batch(self, rho, E, dt, rtol_y, y_floor, rtol_t, max_sub, with_unfinished=False):
        """rho [n][ns]; returns new (rho, E, nsub[, unfinished])."""
        r = np.ascontiguousarray(rho, dtype=np.float64).copy()
        e = np.ascontiguousarray(E, dtype=np.float64).copy()
        d = np.ascontiguousarray(dt, dtype=np.float64)
        ns = np.zeros(len(e), dtype=np.int32)
        un = np.zeros(len(e), dtype=np.int32)
        if self.lib.orc_chem_adaptive_batch(len(e), _ptr(r), _ptr(e), _ptr(d), float(rtol_y), float(y_floor),
                                            float(rtol_t), int(max_sub), _ptr(ns), _ptr(un)):
            raise RuntimeError("orc_chem_adaptive_batch failed")
        return (r, e, ns, un) if with_unfinished else (r, e, ns)
