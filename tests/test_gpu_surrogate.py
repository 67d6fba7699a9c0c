"""The learned chemistry surrogate on the GPU: inference kernels against the
generic MLP path, conservation of accepted cells, fallback cells against the
plain FP64 integrator, and accuracy against the plain integrator.

The network is trained in the module fixture by the library's own kernels
(2^21 jet-flame cells, a short schedule).  Tolerances:
  * fused vs generic forward: both are BF16 MFMA with FP32 accumulation in
    a different order; increments agree to 2e-2 of each species' scale
    out_scale_k at worst and 1e-3 of it at the median;
  * accepted cells conserve mass and H/O/N to 1e-12 relative and
    E + sum_k rho_k R_k F_k to 1e-12 relative (the increment is projected
    onto element balance in FP64);
  * fallback cells: bit-identical to the plain integrator on the same cell;
  * accuracy of accepted cells after this short training: |dT| p99 <= 5 K and
    max <= 5% of T, |dY| <= 1e-2 (the fully trained figures are reported by
    the benchmark).
"""
import numpy as np
import pytest

from conftest import load_package, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    if not gpu_available():
        pytest.skip("no GPU")
    load_package()
    import torch
    from ocfd_amd import chemistry, surrogate, mechanism
    M = mechanism.Mechanism()
    ch = chemistry.ChemBatch(M)
    S = surrogate.Surrogate(ch, seed=0)
    data = S.dataset(1 << 21, seed=1)
    S.configure(data)
    x, y = S.training_arrays(data)
    hist = S.train(x, y, epochs=12, batch=8192, lr=1e-3)
    print("training loss per epoch", [f"{h:.2e}" for h in hist])
    test = S.dataset(1 << 16, seed=99)
    torch.cuda.synchronize()
    return S, ch, M, test


def run(S, test, fallback=True):
    import torch
    rho, E = test["rho0"].clone(), test["E0"].clone()
    flags = torch.zeros(E.shape[0], dtype=torch.int32, device="cuda")
    S.step(rho, E, test["dt"], fallback=fallback, flags=flags)
    torch.cuda.synchronize()
    return rho, E, flags


def inference_vs_generic(S, test):
    """Increments of the inference kernels and of the generic MLP forward
    (features mode 2, scaling, element projection) on the accepted cells, in
    units of each species' out_scale; also the inference increments."""
    rho, E, flags = run(S, test, fallback=False)
    x = S.features(test["rho0"], test["E0"], test["dt"], mode=2)
    z = S.net.forward(x).double().cpu().numpy()
    cfg = S.cfg
    sigma = cfg[24:33]
    P = cfg[S.DY_MAX_INDEX + 1:S.DY_MAX_INDEX + 82].reshape(9, 9)
    dY_gen = (z * sigma) @ P.T
    r0 = test["rho0"].cpu().numpy()
    r = r0.sum(axis=0)
    dY_fused = ((rho.cpu().numpy() - r0) / r[None, :]).T
    acc = flags.cpu().numpy() == 0
    # N2 (index 8) is inert: its increment is zero and its scale a floor
    err = np.abs(dY_fused[acc] - dY_gen[acc])[:, :8] / sigma[None, :8]
    return err, dY_fused


def test_fused_kernel_matches_generic_forward(sur):
    S, ch, M, test = sur
    err, _ = inference_vs_generic(S, test)
    print(f"inference vs generic: max {err.max():.2e}, median {np.median(err):.2e} (in units of out_scale)")
    assert err.max() <= 2e-2 and np.median(err) <= 1e-3


def test_accepted_cells_conserve(sur):
    S, ch, M, test = sur
    rho, E, flags = run(S, test, fallback=False)
    acc = flags.cpu().numpy() == 0
    r0, r1 = test["rho0"].cpu().numpy()[:, acc], rho.cpu().numpy()[:, acc]
    e0, e1 = test["E0"].cpu().numpy()[acc], E.cpu().numpy()[acc]
    m0 = r0.sum(axis=0)
    assert np.max(np.abs(r1.sum(axis=0) - m0) / m0) <= 1e-12
    el0, el1 = (r0 / M.mi[:, None]).T @ M.elements, (r1 / M.mi[:, None]).T @ M.elements
    assert np.max(np.abs(el1 - el0) / (m0[:, None] / 0.001)) <= 1e-12
    hf = (M.coef[:, 0, 5] * 8.314 / M.mi)[:, None]
    t0, t1 = e0 + (r0 * hf).sum(axis=0), e1 + (r1 * hf).sum(axis=0)
    assert np.max(np.abs(t1 - t0) / (np.abs(e0) + np.abs(r0 * hf).sum(axis=0))) <= 1e-12


def test_fallback_cells_equal_plain_integrator(sur):
    import torch
    S, ch, M, test = sur
    # a strict bound on negative mass fractions sends many cells to the fallback
    cfg = S.cfg.copy()
    strict = cfg.copy()
    strict[39] = 0.0
    S.set_config(strict)
    try:
        rho, E, flags = run(S, test, fallback=True)
    finally:
        S.set_config(cfg)
    fb = flags.cpu().numpy() != 0
    assert fb.sum() > 100 and S.fallbacks() == int(fb.sum())
    rp, Ep = test["rho0"].clone(), test["E0"].clone()
    ch.adaptive(rp, Ep, test["dt"])
    torch.cuda.synchronize()
    assert np.array_equal(rho.cpu().numpy()[:, fb], rp.cpu().numpy()[:, fb])
    assert np.array_equal(E.cpu().numpy()[fb], Ep.cpu().numpy()[fb])


def test_accuracy_against_plain_integrator(sur):
    S, ch, M, test = sur
    rho, E, flags = run(S, test, fallback=True)
    Tp = S.temperature(test["rho1"], test["E1"]).cpu().numpy()
    Ts = S.temperature(rho, E).cpu().numpy()
    Yp = (test["rho1"] / test["rho1"].sum(0)).cpu().numpy()
    Ys = (rho / rho.sum(0)).cpu().numpy()
    dT = np.abs(Ts - Tp)
    dY = np.abs(Ys - Yp).max(axis=0)
    print(f"fallbacks {S.fallbacks()} of {len(Tp)}; |dT| max {dT.max():.2f} K p99 {np.percentile(dT, 99):.3f} K; "
          f"max |dT|/T {np.max(dT / Tp):.2e}; |dY| max {dY.max():.2e}")
    assert np.percentile(dT, 99) <= 5.0 and np.max(dT / Tp) <= 0.05 and dY.max() <= 1e-2


def test_domain_check_sends_cells_to_fallback(sur):
    import torch
    S, ch, M, test = sur
    dt = test["dt"].clone()
    dt[:100] = 1e-3  # far outside the trained step sizes
    rho, E = test["rho0"].clone(), test["E0"].clone()
    flags = torch.zeros(E.shape[0], dtype=torch.int32, device="cuda")
    S.step(rho, E, dt, fallback=True, flags=flags)
    f = flags.cpu().numpy()
    assert np.all((f[:100] & 1) != 0)


def test_single_element_streams_pass_the_balance_check(sur):
    """Cells of one stream only (pure air: no H; pure H2/N2 fuel: no O) pass
    the element-balance check: the projected increment conserves every
    element to rounding, and the bound is relative to the cell's total
    element content, so an absent element is not held to zero."""
    import torch
    from ocfd_amd import cases
    S, ch, M, test = sur
    yf, yo = cases.jet_streams(M)
    n = 4096
    rng = np.random.default_rng(5)
    Y = np.where(np.arange(n)[None, :] % 2 == 0, yo[:, None], yf[:, None])
    T = rng.uniform(950.0, 2400.0, n)
    p = 101300.0 * rng.uniform(0.6, 4.5, n)
    rho = np.ascontiguousarray(Y * M.density(Y, T, p)[None, :])
    r = torch.tensor(rho, device="cuda")
    E = torch.tensor(M.energy(rho, T), device="cuda")
    dt = torch.tensor(np.exp(rng.uniform(np.log(2e-8), np.log(5e-7), n)), device="cuda")
    flags = torch.zeros(n, dtype=torch.int32, device="cuda")
    S.step(r, E, dt, fallback=False, flags=flags)
    f = flags.cpu().numpy()
    print("flags of pure-stream cells:", {b: int(np.sum((f & b) != 0)) for b in (1, 2, 4, 8, 16)})
    assert not np.any((f & 2) != 0)


def test_tile_shapes_bitwise_identical_and_autotune(sur):
    """Every tile shape of the network kernel gives the same bits; the
    autotuner leaves the inputs untouched and keeps one of the shapes."""
    import torch
    S, ch, M, test = sur
    outs = []
    for v in range(4):
        S.tile(v)
        rho, E, flags = run(S, test, fallback=False)
        outs.append((rho.cpu().numpy(), E.cpu().numpy(), flags.cpu().numpy()))
    for o in outs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(outs[0], o))
    r0, e0 = test["rho0"].clone(), test["E0"].clone()
    rows, waves, ms = S.autotune(r0, e0, test["dt"], reps=2)
    print("tile shapes (ms):", ms, "chosen", rows, "x", waves)
    assert torch.equal(r0, test["rho0"]) and torch.equal(e0, test["E0"])
    assert (rows, waves) in ((128, 4), (64, 4), (96, 4)) and all(m > 0 for m in ms)


def test_training_gains_over_untrained_network(sur):
    """What the training buys: the same network at its initial weights sends
    most cells to the fallback or errs by far more; after training most cells
    are accepted and the error against the plain integrator is small."""
    import torch
    S, ch, M, test = sur
    trained = S.net.get_params()
    Tp = S.temperature(test["rho1"], test["E1"]).cpu().numpy()
    try:
        S.net.set_params(S.net.init_params(5))       # untrained, same architecture
        rho, E, flags = run(S, test, fallback=False)
        acc_u = flags.cpu().numpy() == 0
        err_u = np.abs(S.temperature(rho, E).cpu().numpy() - Tp)[acc_u]
    finally:
        S.net.set_params(trained)
    rho, E, flags = run(S, test, fallback=False)
    acc_t = flags.cpu().numpy() == 0
    err_t = np.abs(S.temperature(rho, E).cpu().numpy() - Tp)[acc_t]
    p99_u = np.percentile(err_u, 99) if err_u.size else np.inf
    p99_t = np.percentile(err_t, 99)
    print(f"accepted: untrained {acc_u.mean():.3f}, trained {acc_t.mean():.3f}; "
          f"|dT| p99 untrained {p99_u:.1f} K, trained {p99_t:.2f} K")
    assert acc_t.mean() > 0.5
    assert acc_u.mean() < 0.5 * acc_t.mean() or p99_u > 10 * p99_t


def test_calibrated_increment_bound(sur):
    """calibrate() on labelled cells sets dy_max so that, on those same cells,
    every accepted cell is within tol of the plain integrator in every mass
    fraction (the inference is deterministic, so this is exact); exactly
    the cells it reports go to the fallback with flag 16 alone, and there
    they match the plain integrator bit for bit.  On held-out cells the
    bound keeps the error within a few tol."""
    import torch
    S, ch, M, test = sur
    n = test["E0"].shape[0]
    h = n // 2
    part = lambda a, b: {k: v[..., a:b].contiguous() for k, v in test.items()}
    calset, held = part(0, h), part(h, n)
    tol = 2e-4
    saved = S.cfg.copy()
    try:
        cal = S.calibrate(calset, tol=tol)
        print("calibration:", cal)
        assert 0 < cal["dy_max"] < np.inf and cal["sent_to_fallback"] < cal["accepted_unbounded"]

        def err_of(data, fallback):
            rho, E, flags = run(S, data, fallback=fallback)
            Ys = (rho / rho.sum(0)).cpu().numpy()
            Yt = (data["rho1"] / data["rho1"].sum(0)).cpu().numpy()
            return np.abs(Ys - Yt).max(axis=0), flags.cpu().numpy(), rho, E

        err, f, _, _ = err_of(calset, False)
        acc = f == 0
        assert err[acc].max() <= tol
        assert int(np.sum(f == 16)) == cal["sent_to_fallback"]
        # flag-16 cells through the fallback equal the plain integrator
        _, f2, rho, E = err_of(calset, True)
        only16 = torch.from_numpy(f2 == 16).cuda()
        rp, Ep = calset["rho0"].clone(), calset["E0"].clone()
        ch.adaptive(rp, Ep, calset["dt"])
        assert torch.equal(rho[:, only16], rp[:, only16]) and torch.equal(E[only16], Ep[only16])
        err_h, fh, _, _ = err_of(held, False)
        acc_h = fh == 0
        print(f"held-out: accepted {acc_h.mean():.3f}, max error {err_h[acc_h].max():.2e} (tol {tol:.0e})")
        assert acc_h.mean() > 0.05 and err_h[acc_h].max() <= 10 * tol
    finally:
        S.set_config(saved)


def test_fallback_order_is_bitwise_identical(sur):
    """The fallback list taken by decreasing dt (the default) or in the order
    the surrogate kernel appended it gives the same bits."""
    S, ch, M, test = sur
    from ocfd_amd import chemistry
    out = []
    for order in (0, 1):
        ch.set_option(chemistry.OPT_ORDER, order)
        try:
            rho, E, flags = run(S, test, fallback=True)
        finally:
            ch.set_option(chemistry.OPT_ORDER, 1)
        out.append((rho.cpu().numpy(), E.cpu().numpy(), flags.cpu().numpy(), S.fallbacks()))
    assert out[0][3] > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_inference_follows_new_weights(sur):
    """The inference keeps the weights in its own (MFMA fragment) order; after
    set_params it must use the new weights: it matches the generic forward
    with them and differs from the result with the old ones."""
    S, ch, M, test = sur
    p0 = S.net.get_params()
    _, d0 = inference_vs_generic(S, test)
    try:
        S.net.set_params(p0 * np.float32(0.9))
        err, d1 = inference_vs_generic(S, test)
        print(f"new weights: inference vs generic max {err.max():.2e}, median {np.median(err):.2e}")
        assert err.max() <= 2e-2 and np.median(err) <= 1e-3
        assert not np.array_equal(d0, d1)
    finally:
        S.net.set_params(p0)
    _, d2 = inference_vs_generic(S, test)
    assert np.array_equal(d0, d2)
