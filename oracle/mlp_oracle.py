"""FP64 restatement of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

The network the HIP library trains and evaluates: layers W_l [out][in],
b_l [out], GELU in its tanh form after every layer but the last, mean
squared error loss over the true output columns.  ``forward`` and
``gradients`` run in float64; with ``bf16=True`` they round the operands
the GPU rounds (inputs, weights and the activations feeding every GEMM, and
the output-gradient operands of the backward GEMMs) to bfloat16
(round-to-nearest-even), so the only remaining difference to the GPU is its
FP32 accumulation order.  Pinned by central finite differences of the loss
(tests/test_oracle_mlp.py).  No reference counterpart: the surrogate is a
build addition.
"""
import numpy as np

C_GELU = 0.7978845608028654


def to_bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    v = ((u + r) & 0xFFFF0000).astype(np.uint32)
    return v.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def dgelu(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def unflatten(flat, dims):
    """Flat parameter vector (per layer W [out][in] row-major then b) -> lists."""
    Ws, bs, k = [], [], 0
    for l in range(len(dims) - 1):
        din, dout = dims[l], dims[l + 1]
        Ws.append(np.asarray(flat[k:k + din * dout], dtype=np.float64).reshape(dout, din))
        k += din * dout
        bs.append(np.asarray(flat[k:k + dout], dtype=np.float64))
        k += dout
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def forward(flat, dims, x, bf16=False, keep=False):
    Ws, bs = unflatten(flat, dims)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = r(x)
    Zs, As = [], [a]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = a @ r(W).T + b
        Zs.append(z)
        if l < len(Ws) - 1:
            a = r(gelu(z))
            As.append(a)
    return (z, Zs, As) if keep else z


def gradients(flat, dims, x, y, bf16=False):
    """Loss L = mean((net(x) - y)^2) and dL/dparams (flat layout)."""
    Ws, bs = unflatten(flat, dims)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    out, Zs, As = forward(flat, dims, x, bf16, keep=True)
    B, dout = y.shape
    e = out - y
    loss = float(np.mean(e * e))
    dz = 2.0 * e / (B * dout)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dzr = r(dz)
        gW[l] = dzr.T @ As[l]
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dz = (dzr @ r(Ws[l])) * dgelu(Zs[l - 1])
    return loss, flatten(gW, gb)
