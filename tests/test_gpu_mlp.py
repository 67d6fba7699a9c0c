"""GPU MLP kernels (BF16 MFMA, FP32 accumulation) against the FP64 oracle.

Tolerances: GEMMs of bfloat16 operands agree with the float64 product of the
same operands to 1e-5 relative (FP32 accumulation).  Forward outputs and
gradients are compared with the oracle that rounds the same operands to
bfloat16: what remains is the FP32 accumulation order and occasional
activations that round to the neighbouring bfloat16 value, so outputs agree
to 1e-3 of their largest magnitude and each layer's gradient to 2e-2 in
relative Frobenius norm (a transposed operand, a missing GELU' or a wrong
scale gives errors of order 1).
"""
import numpy as np
import pytest

from conftest import load_package, gpu_available
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

DIMS = (12, 256, 256, 256, 256, 256, 9)


@pytest.fixture(scope="module")
def pkg():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_package()


def bf16_dev(a):
    import torch
    return torch.tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).cuda().contiguous()


@pytest.mark.parametrize("ta,tb,splits", [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1), (1, 0, 4), (0, 1, 3)])
def test_gemm_layouts(pkg, ta, tb, splits):
    import torch
    from ocfd_amd import surrogate
    net = surrogate.MLP((32, 64))
    rng = np.random.default_rng(2)
    M, N, K = 128, 192, 320
    A = mo.to_bf16(rng.standard_normal((K, M) if ta else (M, K)))
    B = mo.to_bf16(rng.standard_normal((N, K) if tb else (K, N)))
    C = net.gemm_bf16(bf16_dev(A), bf16_dev(B), ta=bool(ta), tb=bool(tb), splits=splits)
    torch.cuda.synchronize()
    ref = (A.T if ta else A) @ (B.T if tb else B)
    err = np.max(np.abs(C.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"gemm ta={ta} tb={tb} splits={splits}: rel err {err:.2e}")
    assert err < 1e-5


def batch(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, DIMS[0])).astype(np.float32), rng.standard_normal((n, DIMS[-1])).astype(np.float32)


def test_forward_matches_oracle(pkg):
    import torch
    from ocfd_amd import surrogate
    net = surrogate.MLP(DIMS, seed=3)
    flat = net.get_params().astype(np.float64)
    x, _ = batch(1000, 4)
    y = net.forward(torch.tensor(x).cuda()).cpu().numpy()
    ref = mo.forward(flat, DIMS, x.astype(np.float64), bf16=True)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f"forward: max err / max |out| = {err:.2e}")
    assert err < 1e-3


def test_gradients_match_oracle(pkg):
    import torch
    from ocfd_amd import surrogate
    net = surrogate.MLP(DIMS, seed=5)
    flat = net.get_params().astype(np.float64)
    x, y = batch(1000, 6)
    loss, g = net.gradients(torch.tensor(x).cuda(), torch.tensor(y).cuda())
    lo, go = mo.gradients(flat, DIMS, x.astype(np.float64), y.astype(np.float64), bf16=True)
    g = g.cpu().numpy()
    assert abs(loss.item() - lo) <= 1e-4 * lo
    k = 0
    for din, dout in zip(DIMS[:-1], DIMS[1:]):
        for name, size in (("W", din * dout), ("b", dout)):
            a, b = g[k:k + size], go[k:k + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(f"layer {din}->{dout} {name}: rel err {rel:.2e}")
            assert rel < 2e-2
            k += size


def test_training_reduces_loss(pkg):
    """Adam on a fixed regression target (a random teacher network)."""
    import torch
    from ocfd_amd import surrogate
    teacher = surrogate.MLP((12, 64, 9), seed=11)
    xs = torch.tensor(np.random.default_rng(7).standard_normal((8192, 12)).astype(np.float32)).cuda()
    ys = teacher.forward(xs)
    net = surrogate.MLP(DIMS, seed=8)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    hist = []
    for it in range(300):
        i = (it * 1024) % 8192
        net.train_step(xs[i:i + 1024].contiguous(), ys[i:i + 1024].contiguous(), lr=1e-3, loss=loss)
        if it % 50 == 0 or it == 299:
            torch.cuda.synchronize()
            hist.append(loss.item())
    print("loss history", hist)
    assert hist[-1] < 0.1 * hist[0]
