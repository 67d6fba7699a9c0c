"""Pins the FP64 MLP oracle: bfloat16 rounding against torch's conversion,
gradients against central finite differences of the loss."""
import numpy as np
import pytest

from conftest import ROOT  # noqa: F401  (puts the repository root on sys.path)
from oracle import mlp_oracle as mo


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10 ** rng.uniform(-6, 6, 20000),
                        # exact ties between two bfloat16 values
                        (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.to_bf16(x), ref)


def test_gelu_derivative():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.dgelu(z))) < 1e-8
    # GELU(0) = 0, GELU(z) -> z for large z, -> 0 for very negative z
    assert mo.gelu(0.0) == 0.0 and abs(mo.gelu(8.0) - 8.0) < 1e-12 and abs(mo.gelu(-8.0)) < 1e-12


@pytest.mark.parametrize("dims", [(5, 8, 8, 3), (12, 16, 16, 16, 9)])
def test_gradients_match_finite_differences(dims):
    rng = np.random.default_rng(1)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    flat = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((7, dims[0]))
    y = rng.standard_normal((7, dims[-1]))
    loss, g = mo.gradients(flat, dims, x, y)
    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (mo.gradients(flat + e, dims, x, y)[0] - mo.gradients(flat - e, dims, x, y)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    assert abs(loss - np.mean((mo.forward(flat, dims, x) - y) ** 2)) < 1e-14
