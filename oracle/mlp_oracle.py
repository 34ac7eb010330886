"""CPU oracle of the warm-start surrogate network (TEST INFRASTRUCTURE ONLY).

A numpy FP64 restatement of the fully connected network that libgmod.so
trains on the GPU (csrc/nn.hip): SiLU hidden layers, linear output, MSE
loss, with the same BF16 roundings of the GEMM operands (inputs, weights,
activations and back-propagated gradients) so that forward values and
gradients can be compared element by element.  Only tests/ import it.

Pinning: the backward pass here is itself checked against central finite
differences of the FP64 forward pass (tests/test_surrogate_oracle.py), so a
wrong sign, index or transposition in the restated gradient fails there.
"""
import numpy as np


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unflatten(flat, dims):
    Ws, bs, off = [], [], 0
    for i in range(len(dims) - 1):
        nw = dims[i] * dims[i + 1]
        Ws.append(np.asarray(flat[off:off + nw], dtype=np.float64).reshape(dims[i], dims[i + 1]))
        bs.append(np.asarray(flat[off + nw:off + nw + dims[i + 1]], dtype=np.float64))
        off += nw + dims[i + 1]
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def forward(flat, dims, X, emulate_bf16=True):
    """Returns (Y, cache)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unflatten(flat, dims)
    H = [r(X)]
    Z = []
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = H[-1] @ r(W) + b
        Z.append(z)
        if i + 1 < len(Ws):
            H.append(r(silu(z)))
    return Z[-1], (Ws, bs, H, Z, r)


def loss_grad(flat, dims, X, T, emulate_bf16=True):
    """MSE loss mean((Y - T)^2) and its gradient (flat, parameter layout)."""
    Y, (Ws, bs, H, Z, r) = forward(flat, dims, X, emulate_bf16)
    n = Y.size
    loss = float(np.mean((Y - T) ** 2))
    dZ = 2.0 * (Y - T) / n
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for i in range(len(Ws) - 1, -1, -1):
        dZr = r(dZ)
        gb[i] = dZ.sum(axis=0)
        gW[i] = H[i].T @ dZr
        if i > 0:
            dH = dZr @ r(Ws[i]).T
            dZ = dH * dsilu(Z[i - 1])
    return loss, flatten(gW, gb)
