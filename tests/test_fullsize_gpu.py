"""The benchmark configurations at their full sizes, through the C ABI, checked
with size-independent properties on sampled entries:

* DSYRK (lower, N = K = 16384, G ~ U[-1, 1]) then DPOTRF of G G^T + N I,
  tile 1024, FP64: sampled entries of the SYRK against double-double dot
  products (within 2 K eps |G_i||G_j|), info = 0, and sampled entries of
  A - L L^T within 10 N eps ||L_i|| ||L_j||;
* DGEMM 16384^3 with rows of A / columns of B scaled by 2^u, u ~ U[-8, 8],
  C = 1.5 A B - 0.5 C0, in split-BF16 and split-FP8 modes: normwise error of
  sampled entries against the double-double oracle <= 1e-14.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def test_config_dsyrk_then_dpotrf_16384(xk, orc):
    n, nb = 16384, 1024
    L = xk.lib()
    r = np.random.default_rng(2)
    G = np.asfortranarray(r.uniform(-1.0, 1.0, (n, n)))
    S = np.zeros((n, n), order="F")
    L.xkblas_set_param(nb, 0)
    xk.syrk(122, 111, 1.0, G, 0.0, S)
    xk.invalidate()
    rows = r.integers(0, n, 64)
    cols = r.integers(0, n, 64)
    rows, cols = np.maximum(rows, cols), np.minimum(rows, cols)
    Z = np.zeros((1, 1), order="F")
    for i, j in zip(rows, cols):
        gi, gj = np.asfortranarray(G[i:i + 1, :]), np.asfortranarray(G[j:j + 1, :])
        hi, lo = orc.dd_entries(111, 112, 1.0, gi, gj, 0.0, Z, [0], [0])
        bound = 2 * n * EPS * np.linalg.norm(gi) * np.linalg.norm(gj)
        assert abs(S[i, j] - hi[0] - lo[0]) <= bound
    del G
    S[np.diag_indices(n)] += n
    F = S.copy(order="F")
    info = xk.potrf(122, F)
    xk.invalidate()
    assert info == 0
    worst = 0.0
    for i, j in zip(rows, cols):
        li, lj = F[i, :i + 1], F[j, :j + 1]
        res = abs(S[i, j] - li[:j + 1] @ lj) / (np.linalg.norm(li) * np.linalg.norm(lj))
        worst = max(worst, res)
    assert worst <= 10 * n * EPS, worst


@pytest.mark.parametrize("mode", [2, 3], ids=["bf16", "fp8"])
def test_config_split_dgemm_16384(xk, orc, mode):
    n = 16384
    L = xk.lib()
    r = np.random.default_rng(3)
    A = r.uniform(-0.5, 0.5, (n, n))
    A *= np.exp2(r.uniform(-8, 8, n))[:, None]
    A = np.asfortranarray(A)
    B = r.uniform(-0.5, 0.5, (n, n))
    B *= np.exp2(r.uniform(-8, 8, n))[None, :]
    B = np.asfortranarray(B)
    C0 = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    C = C0.copy(order="F")
    L.xkblas_set_modemath(mode)
    try:
        xk.gemm(111, 111, 1.5, A, B, -0.5, C)
    finally:
        L.xkblas_set_modemath(0)
        xk.invalidate()
    rows = r.integers(0, n, 64)
    cols = r.integers(0, n, 64)
    hi, lo = orc.dd_entries(111, 111, 1.5, A, B, -0.5, C0, rows, cols)
    err = (C[rows, cols] - hi) - lo
    assert np.linalg.norm(err) / np.linalg.norm(hi) <= 1e-14


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
def test_config_batched_uneven_full(xk, dtype):
    """The committed batched workload (4096 GEMMs + 1024 TRSMs) in one
    hipGraph-replayed persistent launch; every 64th GEMM and 16th TRSM is
    checked against NumPy in FP64 (FP64: 1e-12 relative, FP32: 1e-5)."""
    import json
    import os
    L = xk.lib()
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_c4_shapes.json")) as f:
        shapes = json.load(f)
    r = np.random.default_rng(11)
    code = 0 if dtype == np.float64 else 1
    probs, keep, checks = [], [], []
    for g, (m, n, k) in enumerate(shapes["gemm_mnk"]):
        A = np.asfortranarray(r.uniform(-0.5, 0.5, (m, k)).astype(dtype))
        B = np.asfortranarray(r.uniform(-0.5, 0.5, (k, n)).astype(dtype))
        C = np.asfortranarray(r.uniform(-0.5, 0.5, (m, n)).astype(dtype))
        dA, dB, dC = xk.DeviceArray(A), xk.DeviceArray(B), xk.DeviceArray(C)
        probs.append(xk.BatchProblem(xk.BATCH_GEMM, code, 111, 111, m, n, k, 1.5, -0.5, dA.ptr, m, dB.ptr, k,
                                     dC.ptr, m))
        keep += [dA, dB, dC]
        if g % 64 == 0:
            ref = 1.5 * A.astype(np.float64) @ B.astype(np.float64) - 0.5 * C.astype(np.float64)
            checks.append((dC, ref))
    for t, n in enumerate(shapes["trsm_n"]):
        Lm = np.tril(r.uniform(-1, 1, (n, n)))
        np.fill_diagonal(Lm, 1.0 + np.abs(Lm).sum(axis=1))
        Lm = np.asfortranarray(Lm.astype(dtype))
        X = np.asfortranarray(r.uniform(-1, 1, (n, 64)).astype(dtype))
        dL, dX = xk.DeviceArray(Lm), xk.DeviceArray(X)
        probs.append(xk.BatchProblem(xk.BATCH_TRSM, code, 111, 111, n, 64, n, 1.0, 0.0, dL.ptr, n, None, 0,
                                     dX.ptr, n))
        keep += [dL, dX]
        if t % 16 == 0:
            checks.append((dX, np.linalg.solve(Lm.astype(np.float64), X.astype(np.float64))))
    arr = (xk.BatchProblem * len(probs))(*probs)
    plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
    assert plan
    assert L.xkblas_batch_use_graph(plan, 1) == 0
    assert L.xkblas_batch_run(plan) == 0
    L.xkblas_sync()
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for d, ref in checks:
        got = d.download().astype(np.float64)
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())
    L.xkblas_batch_use_graph(plan, 0)
    L.xkblas_batch_destroy(plan)
    for d in keep:
        d.free()
