"""CPU oracle of the selector MLP (TEST INFRASTRUCTURE ONLY).

Same parameter layout as the GPU kernels (pava_amd/csrc/mlp.hip): W1 [H x I],
b1, W2 [H x H], b2, W3 [O x H], b3; ReLU hidden layers; masked MSE loss
sum M (Y - T)^2 / sum M. Computed in float64 numpy. Pinned by central finite
differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def unpack(p, I, H, O):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for rows, cols in ((H, I), (H, 0), (H, H), (H, 0), (O, H), (O, 0)):
        n = rows * max(cols, 1)
        a = p[o:o + n]
        out.append(a.reshape(rows, cols) if cols else a)
        o += n
    assert o == p.size
    return out


def forward(p, X, I, H, O):
    W1, b1, W2, b2, W3, b3 = unpack(p, I, H, O)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    Y = A2 @ W3.T + b3
    return Y, (X, Z1, A1, Z2, A2)


def loss(p, X, T, M, I, H, O):
    Y, _ = forward(p, X, I, H, O)
    M = np.asarray(M, np.float64)
    return float((M * (Y - T) ** 2).sum() / M.sum())


def loss_grad(p, X, T, M, I, H, O):
    W1, b1, W2, b2, W3, b3 = unpack(p, I, H, O)
    Y, (X, Z1, A1, Z2, A2) = forward(p, X, I, H, O)
    M = np.asarray(M, np.float64)
    T = np.asarray(T, np.float64)
    w = M.sum()
    L = float((M * (Y - T) ** 2).sum() / w)
    dY = 2 * M * (Y - T) / w
    gW3 = dY.T @ A2
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3) * (Z2 > 0)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g
