"""The tile-cost MLP trained by the library's own HIP kernels: forward and
full-batch gradient against the CPU restatement (oracle/mlp.py), and Adam
training that fits a smooth function."""
import ctypes

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_forward_and_gradient_match_cpu(xk):
    L = xk.lib()
    r = np.random.default_rng(1)
    n_in, n = 12, 1000
    h = L.xkblas_mlp_create(n_in, 5)
    P = L.xkblas_mlp_num_params(h)
    _, Pc = mlp.layout(n_in)
    assert P == Pc
    params = np.zeros(P, dtype=np.float32)
    L.xkblas_mlp_get_params(h, _p(params))
    X = r.normal(0, 1, (n, n_in)).astype(np.float32)
    T = r.normal(0, 1, n).astype(np.float32)
    y = np.zeros(n, dtype=np.float32)
    assert L.xkblas_mlp_forward(h, _p(X), n, _p(y)) == 0
    y_ref, _ = mlp.forward(params.astype(np.float64), X.astype(np.float64))
    np.testing.assert_allclose(y, y_ref, rtol=1e-4, atol=1e-4)
    g = np.zeros(P, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    assert L.xkblas_mlp_gradient(h, _p(X), _p(T), n, _p(g), _p(loss)) == 0
    g_ref, l_ref = mlp.gradient(params.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(loss[0] - l_ref) <= 1e-4 * max(1.0, l_ref)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert err <= 1e-4, err
    L.xkblas_mlp_destroy(h)


def test_training_fits_smooth_target(xk):
    L = xk.lib()
    r = np.random.default_rng(2)
    n_in, n = 4, 4096
    X = r.uniform(-1, 1, (n, n_in)).astype(np.float32)
    T = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3] ** 2).astype(np.float32)
    h = L.xkblas_mlp_create(n_in, 3)
    params = np.zeros(L.xkblas_mlp_num_params(h), dtype=np.float32)
    L.xkblas_mlp_get_params(h, _p(params))
    l0 = mlp.loss(params.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    last = L.xkblas_mlp_train(h, _p(X), _p(T), n, 300, 256, 3e-3)
    L.xkblas_mlp_get_params(h, _p(params))
    l1 = mlp.loss(params.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert l1 < 0.05 * l0, (l0, l1, last)
    L.xkblas_mlp_destroy(h)
