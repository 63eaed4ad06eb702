"""Pins the numpy MLP oracle by central finite differences (CPU)."""
from __future__ import annotations

import numpy as np

from oracle import mlp


def test_mlp_gradient_finite_differences():
    K = 12
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp.num_params(K)) * 0.2
    X = rng.standard_normal((9, 12))
    T = rng.standard_normal((9, K))
    M = (rng.random((9, K)) < 0.8).astype(np.float64)
    _, g = mlp.backward(p, X, T, M, K)
    idx = rng.choice(len(p), 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, X, T, M, K) - mlp.loss(p - e, X, T, M, K)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_masked_entries_do_not_contribute():
    K = 12
    rng = np.random.default_rng(1)
    p = rng.standard_normal(mlp.num_params(K)) * 0.2
    X = rng.standard_normal((4, 12))
    T = rng.standard_normal((4, K))
    M = np.ones((4, K))
    M[:, 5] = 0.0
    T2 = T.copy()
    T2[:, 5] += 100.0
    assert mlp.loss(p, X, T, M, K) == mlp.loss(p, X, T2, M, K)
