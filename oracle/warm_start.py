"""FP64 restatement of the learned warm-start network (test infrastructure
only: the product's network runs on the GPU, in silk_physics_amd/csrc/hip/
warm_start.hip).  No counterpart in the reference.

Network 16 -> 64 -> 64 -> 1 with SiLU after the hidden layers; loss = mean
squared error.  Parameters are flat, in the order W1 [64][16], b1 [64],
W2 [64][64], b2 [64], W3 [64], b3.  With ``emulate=True`` the BF16
roundings of the device kernels are applied (weights, inputs, activations and
the backward operands that enter BF16 products) so that the remaining
difference is FP32 against FP64 accumulation.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 64
W1, B1 = 0, N_H * N_IN
W2 = B1 + N_H
B2 = W2 + N_H * N_H
W3 = B2 + N_H
B3 = W3 + N_H
N_PARAMS = B3 + 1


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(N_H, N_IN), p[B1:W2], p[W2:B2].reshape(N_H, N_H),
            p[B2:W3], p[W3:B3], p[B3])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, emulate=False):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    X = r(X)
    z1 = X @ r(w1).T + b1
    h1 = r(_silu(z1))
    z2 = h1 @ r(w2).T + b2
    h2 = r(_silu(z2))
    y = h2 @ r(w3) + b3
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(p, X, t, emulate=False):
    """Mean squared error of the predictions and its gradient (N_PARAMS)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (Xr, z1, h1, z2, h2) = forward(p, X, emulate)
    n = y.size
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    dyr = r(dy)
    dz2 = (dyr[:, None] * r(w3)[None, :]) * _dsilu(z2)
    dz2r = r(dz2)
    dz1 = (dz2r @ r(w2)) * _dsilu(z1)
    g = np.zeros(N_PARAMS)
    g[W3:B3] = dyr @ h2
    g[B3] = dy.sum()
    g[W2:B2] = (dz2r.T @ h1).reshape(-1)
    g[B2:W3] = dz2.sum(axis=0)
    g[W1:B1] = (r(dz1).T @ Xr).reshape(-1)
    g[B1:W2] = dz1.sum(axis=0)
    return loss, g, y
