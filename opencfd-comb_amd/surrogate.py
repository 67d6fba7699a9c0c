"""Learned chemistry surrogate: the MLP and its GPU training loop.

``MLP`` wraps the library's BF16-MFMA network (forward, gradients, Adam
steps; FP32 master weights on the device).  Arrays are torch float32 CUDA
tensors (device memory plumbing); without the HIP library every call raises.
The surrogate is a build addition with no reference counterpart.
"""
import ctypes

import numpy as np

from . import load_library, OcfdError


def _vp(p):
    return ctypes.c_void_p(p)


class MLP:
    """dims[0] -> ... -> dims[-1], GELU between layers."""

    def __init__(self, dims, seed=0, lib=None):
        self.lib = lib or load_library()
        self.dims = [int(d) for d in dims]
        arr = (ctypes.c_int32 * len(self.dims))(*self.dims)
        self._check(self.lib.ocfd_mlp_create(len(self.dims) - 1, arr))
        self.n_params = int(self.lib.ocfd_mlp_param_count())
        self.set_params(self.init_params(seed))

    def init_params(self, seed=0):
        """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            s = 1.0 / np.sqrt(din)
            parts.append(rng.uniform(-s, s, dout * din))
            parts.append(rng.uniform(-s, s, dout))
        return np.concatenate(parts).astype(np.float32)

    def _check(self, rc):
        if rc != 0:
            raise OcfdError(self.lib.ocfd_mlp_last_error().decode())

    def _stream(self):
        import torch
        self._check(self.lib.ocfd_mlp_set_stream(_vp(torch.cuda.current_stream().cuda_stream)))

    @staticmethod
    def _dev(t, shape, name):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        return _vp(t.data_ptr())

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        self._check(self.lib.ocfd_mlp_set_params(flat.ctypes.data_as(ctypes.c_void_p)))

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.lib.ocfd_mlp_get_params(out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def forward(self, x):
        import torch
        B = x.shape[0]
        y = torch.empty((B, self.dims[-1]), dtype=torch.float32, device=x.device)
        self._stream()
        self._check(self.lib.ocfd_mlp_forward(B, self._dev(x, (B, self.dims[0]), "x"), _vp(y.data_ptr())))
        return y

    def gradients(self, x, y):
        """(loss, flat gradient) as torch tensors on the device."""
        import torch
        B = x.shape[0]
        g = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        self._stream()
        self._check(self.lib.ocfd_mlp_gradients(B, self._dev(x, (B, self.dims[0]), "x"),
                                                self._dev(y, (B, self.dims[-1]), "y"), _vp(g.data_ptr()),
                                                _vp(loss.data_ptr())))
        return loss, g

    def train_step(self, x, y, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, loss=None):
        """One Adam step on (x, y); the loss lands in ``loss`` (device) if given."""
        B = x.shape[0]
        self._stream()
        self._check(self.lib.ocfd_mlp_train_step(B, self._dev(x, (B, self.dims[0]), "x"),
                                                 self._dev(y, (B, self.dims[-1]), "y"), float(lr), float(beta1),
                                                 float(beta2), float(eps),
                                                 _vp(loss.data_ptr()) if loss is not None else None))

    def gemm_bf16(self, A, B, ta=False, tb=False, splits=1):
        """C = op(A) op(B) with bf16 device operands (tests)."""
        import torch
        M = A.shape[1] if ta else A.shape[0]
        K = A.shape[0] if ta else A.shape[1]
        N = B.shape[0] if tb else B.shape[1]
        C = torch.empty((M, N), dtype=torch.float32, device=A.device)
        self._stream()
        self._check(self.lib.ocfd_mlp_gemm_bf16(M, N, K, int(ta), int(tb), _vp(A.data_ptr()), A.shape[1],
                                                _vp(B.data_ptr()), B.shape[1], _vp(C.data_ptr()), N, splits))
        return C


# ---------------------------------------------------------------------------
# Chemistry surrogate: 12 -> 256 x 5 -> 9 network from cell state to the
# mass-fraction increments of one chemistry step.
# ---------------------------------------------------------------------------
DIMS = (12, 256, 256, 256, 256, 256, 9)
CFG_LEN = 152
N_SPEC = 9


def element_matrix(mech):
    """Cel[e][k]: moles of element e (H, O, N) per kg of species k."""
    return (mech.elements / mech.mi[:, None]).T.copy()


def trust_bound(m, err, tol):
    """Largest increment bound t such that every calibration cell with
    m <= t has err <= tol (m: the cell's largest |dY_k| as the surrogate
    predicts it, err: its error against the plain integrator).  The bound
    sits midway between the last cell kept and the first cell whose error
    exceeds tol, so rounding of m cannot let that cell through; with no such
    cell it is just above the largest m seen (cells beyond it were never
    measured).  0 when the smallest increment already errs by more than tol."""
    m = np.asarray(m, dtype=np.float64)
    err = np.asarray(err, dtype=np.float64)
    if m.size == 0:
        return 0.0
    o = np.argsort(m, kind="stable")
    ms, es = m[o], err[o]
    bad = np.nonzero(~(es <= tol))[0]
    if bad.size == 0:
        return float(ms[-1] * (1.0 + 1e-6))
    mv = ms[bad[0]]
    kept = ms[ms < mv]
    return 0.5 * (float(kept.max()) + float(mv)) if kept.size else 0.0


def projection(cel, sigma):
    """P with P dY the closest (in the sigma-scaled norm) increment that
    conserves every element: P = I - D C^T (C D C^T)^+ C, D = diag(sigma^2)."""
    D = np.diag(np.asarray(sigma, dtype=np.float64) ** 2)
    C = np.asarray(cel, dtype=np.float64)
    return np.eye(C.shape[1]) - D @ C.T @ np.linalg.pinv(C @ D @ C.T) @ C


class Surrogate:
    """Learned chemistry step with element-balance / bounds checks and the
    plain FP64 integrator as fallback (and as the source of its training
    labels and the measure of its error)."""

    def __init__(self, chem, dims=DIMS, seed=0):
        self.chem = chem
        self.lib = chem.lib
        self.mech = chem.mech
        self.net = MLP(dims, seed=seed, lib=self.lib)
        self.cfg = None

    def _check(self, rc):
        if rc != 0:
            raise OcfdError(self.lib.ocfd_chem_last_error().decode())

    # -- data ----------------------------------------------------------------
    def dataset(self, n, seed=0, controller=None, cold_fraction=0.125, T_cold=(280.0, 900.0)):
        """n jet-flame cells along their ignition trajectories (progress by
        the plain integrator) and their state after one chemistry step dt by
        the plain adaptive FP64 integrator.  A fraction cold_fraction of the
        cells is drawn with T in T_cold (the planar jet's 300 K fuel stream and
        its cold mixing layer, where the chemistry is frozen), the rest with T
        in 900-2500 K; the cells are shuffled so any prefix samples both.
        Device tensors."""
        import torch
        from . import cases
        from .chemistry import RTOL_Y, Y_FLOOR, RTOL_T
        c = controller or (RTOL_Y, Y_FLOOR, RTOL_T)
        n_cold = int(round(n * cold_fraction))
        s = cases.jet_chem_states(n - n_cold, seed=seed, mech=self.mech)
        if n_cold:
            sc = cases.jet_chem_states(n_cold, seed=seed + 1000, mech=self.mech, T_range=T_cold)
            perm = np.random.default_rng(seed).permutation(n)
            s = {k: np.concatenate([s[k], sc[k]], axis=-1)[..., perm] for k in ("rho", "E", "dt", "tau")}
        dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda").contiguous()
        rho, E, tau, dt = dev(s["rho"]), dev(s["E"]), dev(s["tau"]), dev(s["dt"])
        self.chem.adaptive(rho, E, tau, *c, max_sub=100000)
        if self.chem.unfinished():
            raise RuntimeError("dataset: progress integration left cells unfinished")
        rho0, E0 = rho.clone(), E.clone()
        self.chem.adaptive(rho, E, dt, *c, max_sub=100000)
        if self.chem.unfinished():
            raise RuntimeError("dataset: label integration left cells unfinished")
        return dict(rho0=rho0, E0=E0, dt=dt, rho1=rho, E1=E)

    def features(self, rho, E, dt=None, mode=1):
        import torch
        n = E.shape[0]
        out = torch.empty((n,) if mode == 0 else (n, 12), dtype=torch.float64 if mode < 2 else torch.float32,
                          device=E.device)
        self.chem._args(rho, E, dt if dt is not None else E)
        self._check(self.lib.ocfd_chem_cell_features(n, _vp(rho.data_ptr()), _vp(E.data_ptr()),
                                                     _vp(dt.data_ptr()) if dt is not None else None, mode,
                                                     _vp(out.data_ptr())))
        return out

    def temperature(self, rho, E):
        return self.features(rho, E, None, mode=0)

    def targets(self, rho0, rho1):
        import torch
        n = rho0.shape[1]
        y = torch.empty((n, N_SPEC), dtype=torch.float32, device=rho0.device)
        self._check(self.lib.ocfd_chem_surrogate_targets(n, _vp(rho0.data_ptr()), _vp(rho1.data_ptr()),
                                                         _vp(y.data_ptr())))
        return y

    # -- configuration ---------------------------------------------------------
    def configure(self, data, sample=1 << 18, y_neg_tol=1e-5, elem_tol=1e-9, T_out=(250.0, None), dy_max=np.inf):
        """Normalisation and acceptance domain from the training data
        (dy_max: no increment bound until calibrate sets one)."""
        n = data["E0"].shape[0]
        k = min(n, sample)
        fall = self.features(data["rho0"], data["E0"], data["dt"], mode=1)
        # the acceptance domain spans every training cell (the normalisation
        # uses a sample): cells just beyond a sample's extremes are states the
        # network was trained on
        fmin, fmax = fall.amin(0).cpu().numpy(), fall.amax(0).cpu().numpy()
        f = fall[:k].cpu().numpy()
        del fall
        r0 = data["rho0"][:, :k].cpu().numpy()
        r1 = data["rho1"][:, :k].cpu().numpy()
        dY = ((r1 - r0) / r0.sum(axis=0)[None, :]).T
        shift = f.mean(axis=0)
        scale = 1.0 / np.maximum(f.std(axis=0), 1e-12)
        sigma = np.maximum(dY.std(axis=0), 1e-12)
        cel = element_matrix(self.mech)
        P = projection(cel, sigma)
        cfg = np.concatenate([
            shift, scale, sigma,
            [fmin[0], fmax[0], np.exp(fmin[1]), np.exp(fmax[1]), np.exp(fmin[11]), np.exp(fmax[11])],
            [y_neg_tol, elem_tol, T_out[0], T_out[1] if T_out[1] is not None else 1.1 * fmax[0], dy_max],
            P.ravel(), cel.ravel()])
        assert cfg.size == CFG_LEN
        self.set_config(cfg)
        return cfg

    T_LO_INDEX = 12 + 12 + N_SPEC          # lower end of the trained temperature domain
    DY_MAX_INDEX = 12 + 12 + N_SPEC + 6 + 4

    def calibrate(self, data, tol=5e-4):
        """Sets the increment bound dy_max from labelled cells the network
        was not trained on: the inference runs on them with no bound and
        no fallback, and among the cells it accepts, dy_max is the largest
        increment size below which every cell's mass fractions (all species)
        are within tol of the plain integrator's (trust_bound).  Returns the
        bound and how many of those cells it sends to the fallback."""
        import torch
        if self.cfg is None:
            raise RuntimeError("calibrate: configure the surrogate first")
        saved = self.cfg.copy()
        cfg = saved.copy()
        cfg[self.DY_MAX_INDEX] = np.inf
        self.set_config(cfg)
        rho, E = data["rho0"].clone(), data["E0"].clone()
        flags = torch.zeros(E.shape[0], dtype=torch.int32, device=E.device)
        try:
            self.step(rho, E, data["dt"], fallback=False, flags=flags)
        except Exception:
            self.set_config(saved)
            raise
        r0 = data["rho0"].sum(0)
        m = ((rho - data["rho0"]) / r0).abs().amax(0)
        Ys = rho / rho.sum(0)
        Yt = data["rho1"] / data["rho1"].sum(0)
        err = (Ys - Yt).abs().amax(0)
        acc = flags == 0
        m, err = m[acc].cpu().numpy(), err[acc].cpu().numpy()
        bound = trust_bound(m, err, tol)
        cfg[self.DY_MAX_INDEX] = bound
        self.set_config(cfg)
        return dict(dy_max=bound, tol=tol, cells=int(E.shape[0]), accepted_unbounded=int(m.size),
                    sent_to_fallback=int(np.sum(~(m <= bound))), err_max_unbounded=float(err.max()) if err.size else 0.0)

    def set_config(self, cfg):
        cfg = np.ascontiguousarray(cfg, dtype=np.float64)
        self._check(self.lib.ocfd_chem_surrogate_set(cfg.ctypes.data_as(ctypes.c_void_p), cfg.size))
        self.cfg = cfg

    # -- training --------------------------------------------------------------
    def training_arrays(self, data):
        x = self.features(data["rho0"], data["E0"], data["dt"], mode=2)
        y = self.targets(data["rho0"], data["rho1"])
        return x, y

    def train(self, x, y, epochs=10, batch=8192, lr=1e-3, lr_final=2e-5, seed=0, time_budget=None, log=None):
        """Adam with a cosine learning-rate decay; returns the loss history
        (one mean per epoch)."""
        import math
        import time
        import torch
        n = x.shape[0]
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        acc = torch.zeros(1, dtype=torch.float64, device="cuda")
        hist, step, t0 = [], 0, time.time()
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=g)
            acc.zero_()
            for s in range(steps_per_epoch):
                idx = perm[s * batch:(s + 1) * batch]
                xb, yb = x[idx].contiguous(), y[idx].contiguous()
                a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / max(1, total - 1)))
                self.net.train_step(xb, yb, lr=a, loss=loss)
                acc += loss
                step += 1
            torch.cuda.synchronize()
            hist.append(acc.item() / steps_per_epoch)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e} ({time.time() - t0:.1f} s)")
            if time_budget is not None and time.time() - t0 > time_budget:
                break
        return hist

    # -- inference ---------------------------------------------------------------
    def step(self, rho, E, dt, fallback=True, controller=None, max_sub=4000, flags=None):
        """One chemistry step of every cell in place; fills the per-cell flags
        if given (0 surrogate, else the failed checks: 1 domain, 2 element
        balance, 4 negative mass fraction, 8 temperature bounds, 16 an
        increment above the calibrated bound)."""
        from .chemistry import RTOL_Y, Y_FLOOR, RTOL_T
        c = controller or (RTOL_Y, Y_FLOOR, RTOL_T)
        n = self.chem._args(rho, E, dt)
        self._check(self.lib.ocfd_chem_surrogate_batch(n, _vp(rho.data_ptr()), _vp(E.data_ptr()), _vp(dt.data_ptr()),
                                                       1 if fallback else 0, float(c[0]), float(c[1]), float(c[2]),
                                                       int(max_sub),
                                                       _vp(flags.data_ptr()) if flags is not None else None))
        return flags

    def autotune(self, rho, E, dt, reps=3):
        """Time every tile shape of the network kernel on scratch copies of the
        batch and keep the fastest (bit-identical results).  Returns
        (rows, waves, ms per shape)."""
        n = self.chem._args(rho, E, dt)
        ms = (ctypes.c_float * 4)()
        self._check(self.lib.ocfd_chem_surrogate_autotune(n, _vp(rho.data_ptr()), _vp(E.data_ptr()),
                                                          _vp(dt.data_ptr()), int(reps), ms))
        rows, waves = self.tile()
        return rows, waves, list(ms)

    def tile(self, variant=-1):
        """(cells per workgroup, waves) of the shape in use; variant >= 0 forces one."""
        r, w = ctypes.c_int32(0), ctypes.c_int32(0)
        v = self.lib.ocfd_chem_surrogate_tile(int(variant), ctypes.byref(r), ctypes.byref(w))
        if v < 0:
            self._check(1)
        return r.value, w.value

    def fallbacks(self):
        v = self.lib.ocfd_chem_surrogate_fallbacks()
        if v < 0:
            raise OcfdError(self.lib.ocfd_chem_last_error().decode())
        return int(v)
