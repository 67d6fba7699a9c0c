"""Batched finite-rate chemistry on the GPU.

``ChemBatch`` mirrors the reference's per-cell 0-D integrator
``chemical_timeAdv(di, E, dt)`` (CHEM_reactions.f90:374-390) and the loop of
``Chemical_time_advance`` (OCFD_Time_Adv.f90:169-196) that applies it
``Chemical_sub_step`` times per cell, over a whole batch of cells at once.
The arrays are torch float64 CUDA tensors (device memory plumbing only):
``rho`` [n_spec, n] species-major partial densities (kg/m3), ``E`` [n]
(J/m3) and ``dt`` [n] (s); they are advanced in place by the HIP kernels.
There is no host path: without the HIP library every call raises.
"""
import ctypes

import numpy as np

from . import load_library, OcfdError
from .mechanism import Mechanism

MODES = {1: "Euler", 2: "RK2", 3: "RK3", -1: "point-implicit Euler"}
OPT_COOP_BUDGET, OPT_FALLBACK_SCHEDULER, OPT_COOP_DIRECT_MAX, OPT_ORDER = 1, 2, 3, 4

# adaptive controller defaults: on jet-mixture cells along their ignition
# trajectories, one step of dt in [1e-8, 1e-6] s lands within 0.25 K and 1e-3
# in mass fraction of a 2000-substep point-implicit solution
RTOL_Y, Y_FLOOR, RTOL_T = 0.05, 1e-6, 0.01


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data)


class ChemBatch:
    """Device-side chemistry for a batch of independent cells."""

    def __init__(self, mech=None, lib=None):
        self.lib = lib or load_library()
        self.mech = mech or Mechanism()
        m = self.mech
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._keep = arrs = dict(
            af=i32(m.af_type), tb=i32(m.third_body), fc=f64(m.fc_troe), arr=f64(m.arrhenius),
            vf=i32(m.vf), vr=i32(m.vr), eff=f64(m.eff), tct=f64(np.append(m.tct, 0.0)), mi=f64(m.mi),
            coef=f64(m.coef))
        self._check(self.lib.ocfd_chem_set_mechanism(
            m.n_spec, m.n_reac, m.rereac, _vp(arrs["af"]), _vp(arrs["tb"]), _vp(arrs["fc"]), _vp(arrs["arr"]),
            _vp(arrs["vf"]), _vp(arrs["vr"]), _vp(arrs["eff"]), m.n_tct, _vp(arrs["tct"]), float(m.tct_max),
            _vp(arrs["mi"]), _vp(arrs["coef"])))
        if m.rereac != 0:
            arrs["rev"] = f64(m.reverse)
            self._check(self.lib.ocfd_chem_set_reverse_rates(_vp(arrs["rev"])))
        self.n_spec = m.n_spec

    def _check(self, rc):
        if rc != 0:
            raise OcfdError(self.lib.ocfd_chem_last_error().decode())

    def _args(self, rho, E, dt):
        import torch
        n = E.shape[0]
        for name, t, shape in (("rho", rho, (self.n_spec, n)), ("E", E, (n,)), ("dt", dt, (n,))):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
                raise TypeError(f"{name} must be a float64 CUDA tensor")
            if tuple(t.shape) != shape or not t.is_contiguous():
                raise ValueError(f"{name} must be contiguous with shape {shape}, got {tuple(t.shape)}")
        # run on torch's current stream so tensor ordering and torch events hold
        self._check(self.lib.ocfd_chem_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return n

    def chemical_timeAdv(self, rho, E, dt, mode=-1, nsub=1):
        """nsub fixed substeps of dt/nsub per cell with the CHEM_TimeAdv
        integrator ``mode`` (1, 2, 3 or -1)."""
        if mode not in MODES:
            raise ValueError("mode (CHEM_TimeAdv) must be 1, 2, 3 or -1")
        n = self._args(rho, E, dt)
        self._check(self.lib.ocfd_chem_timeadv_batch(n, rho.data_ptr(), E.data_ptr(), dt.data_ptr(), mode, nsub))

    def adaptive(self, rho, E, dt, rtol_y=RTOL_Y, y_floor=Y_FLOOR, rtol_t=RTOL_T, max_sub=4000, scheduler=1,
                 nsub_out=None):
        """Point-implicit Euler with per-cell adaptive substeps (controller
        documented at ocfd_chem_adaptive_batch).  scheduler 1 = work queue,
        2 = work queue with the cooperative kernel for long-running cells,
        0 = one cell per thread (bit-identical results).  Returns the int32
        tensor of accepted substeps per cell."""
        import torch
        n = self._args(rho, E, dt)
        if nsub_out is None:
            nsub_out = torch.empty(n, dtype=torch.int32, device=E.device)
        elif not (nsub_out.is_cuda and nsub_out.dtype == torch.int32 and tuple(nsub_out.shape) == (n,)
                  and nsub_out.is_contiguous()):
            raise ValueError("nsub_out must be a contiguous int32 CUDA tensor of shape (n,)")
        self._check(self.lib.ocfd_chem_adaptive_batch(n, rho.data_ptr(), E.data_ptr(), dt.data_ptr(), float(rtol_y),
                                                      float(y_floor), float(rtol_t), int(max_sub),
                                                      nsub_out.data_ptr(), int(scheduler)))
        return nsub_out

    def set_option(self, key, value):
        """OPT_COOP_BUDGET (1), OPT_FALLBACK_SCHEDULER (2), OPT_COOP_DIRECT_MAX (3)
        or OPT_ORDER (4: 1 longest dt first, 0 index order)."""
        self._check(self.lib.ocfd_chem_set_option(int(key), int(value)))

    def counters(self):
        """Work counters of the last call: dict(unfinished, fallbacks, cooperative)."""
        out = (ctypes.c_int64 * 5)()
        self._check(self.lib.ocfd_chem_counters(out, 5))
        return dict(unfinished=int(out[1]), fallbacks=int(out[2]), cooperative=int(out[3]))

    def unfinished(self):
        """Cells the last ``adaptive`` call left short of dt (synchronises)."""
        v = self.lib.ocfd_chem_unfinished()
        if v < 0:
            raise OcfdError(self.lib.ocfd_chem_last_error().decode())
        return int(v)

    def synchronize(self):
        self._check(self.lib.ocfd_chem_synchronize())
