"""Pins of the CPU MLP restatement (no GPU): hand-derived gradients equal
central finite differences of the loss in float64."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    r = np.random.default_rng(0)
    n_in, n = 7, 40
    _, P = mlp.layout(n_in)
    p = r.normal(0, 0.3, P)
    X = r.normal(0, 1, (n, n_in))
    T = r.normal(0, 1, n)
    g, _ = mlp.gradient(p, X, T)
    idx = r.choice(P, 80, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(P)
        e[i] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_layout_size():
    _, P = mlp.layout(12)
    assert P == 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1
