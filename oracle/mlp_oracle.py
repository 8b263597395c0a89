"""NumPy restatement of the surrogate MLP (forward, MSE loss, backward).

TEST INFRASTRUCTURE ONLY: the checker for the engine's MLP kernels
(spicier_amd/csrc/surrogate.hip).  No counterpart exists in the reference.

Network: n_in -> H -> H -> n_out, SiLU after the two hidden layers, weights
packed as W1[n_in][H], b1[H], W2[H][H], b2[H], W3[H][n_out], b3[n_out].
With ``bf16=True`` every matmul operand is rounded to BF16 (round to nearest
even) the way the engine stages MFMA operands, and products are accumulated
in float64 (the engine accumulates in FP32); with ``bf16=False`` everything is
float64, the mode the finite-difference check of this file uses.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w, n_in, H, n_out):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    out = []
    for shape in ((n_in, H), (H,), (H, H), (H,), (H, n_out), (n_out,)):
        size = int(np.prod(shape))
        out.append(w[o:o + size].reshape(shape))
        o += size
    assert o == w.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, X, n_in, H, n_out, bf=True):
    W1, b1, W2, b2, W3, b3 = unpack(w, n_in, H, n_out)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    Z1 = r(X) @ r(W1) + b1
    H1 = silu(Z1)
    Z2 = r(H1) @ r(W2) + b2
    H2 = silu(Z2)
    Y = r(H2) @ r(W3) + b3
    return Y, (X, Z1, H1, Z2, H2)


def loss_and_grad(w, X, T, n_in, H, n_out, bf=True):
    W1, b1, W2, b2, W3, b3 = unpack(w, n_in, H, n_out)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (X, Z1, H1, Z2, H2) = forward(w, X, n_in, H, n_out, bf)
    T = np.asarray(T, dtype=np.float64)
    cnt = Y.size
    d = Y - T
    loss = float((d * d).sum() / cnt)
    dY = 2.0 * d / cnt
    if bf:
        dY = dY.astype(np.float32)
    gW3 = r(H2).T @ r(dY)
    gb3 = np.asarray(dY, dtype=np.float64).sum(0)
    dH2 = r(dY) @ r(W3).T
    dZ2 = dH2 * dsilu(Z2)
    if bf:
        dZ2 = dZ2.astype(np.float32)
    gW2 = r(H1).T @ r(dZ2)
    gb2 = np.asarray(dZ2, dtype=np.float64).sum(0)
    dH1 = r(dZ2) @ r(W2).T
    dZ1 = dH1 * dsilu(Z1)
    if bf:
        dZ1 = dZ1.astype(np.float32)
    gW1 = r(X).T @ r(dZ1)
    gb1 = np.asarray(dZ1, dtype=np.float64).sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
