"""numpy reference of the selector MLP — TEST INFRASTRUCTURE ONLY.

Same parameter layout as the library (flat FP32: W1[12x64] b1[64] W2[64x64]
b2[64] W3[64xK] b3[K], W input-major) and the same loss: masked mean squared
error sum(M*(out-T)^2)/max(1, sum M).  Evaluated in float64 so it can serve as
the gradient oracle of the HIP forward/backward kernels; itself pinned by
central finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np

F, H = 12, 64


def unpack(p, K):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * K].reshape(H, K); o += H * K
    b3 = p[o:o + K]; o += K
    assert o == len(p)
    return W1, b1, W2, b2, W3, b3


def num_params(K):
    return F * H + H + H * H + H + H * K + K


def forward(p, X, K):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h1, h2, h2 @ W3 + b3


def loss(p, X, T, M, K):
    _, _, out = forward(p, X, K)
    M = np.asarray(M, dtype=np.float64)
    return float(np.sum(M * (out - T) ** 2) / max(1.0, M.sum()))


def backward(p, X, T, M, K):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    X = np.asarray(X, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    h1, h2, out = forward(p, X, K)
    cnt = max(1.0, M.sum())
    d = 2.0 * M * (out - T) / cnt
    gW3 = h2.T @ d
    gb3 = d.sum(axis=0)
    dh2 = (d @ W3.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = X.T @ dh1
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.sum(M * (out - T) ** 2) / cnt), g
