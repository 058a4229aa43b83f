"""Batched uneven GEMM + TRSM problems, FP64 and FP32, in one persistent
launch (xkblas_batch_*), through the C ABI.

Bars: FP64 GEMM within the reference GEMM bound (3 eps normwise) against the
oracle's tiled GEMM; FP64 TRSM within the reference TRSM bound (<= 10);
FP32 results within 1e-5 relative of an FP64 evaluation.  hipGraph replays
are checked against the eager run after every replay."""
import ctypes

import numpy as np
import pytest

from helpers import rng

pytestmark = pytest.mark.gpu


def make_batch(xk, r, ngemm, ntrsm, maxdim=200, maxtri=300):
    probs, keep, checks = [], [], []
    for g in range(ngemm):
        m, n, k = (int(x) for x in r.integers(1, maxdim, 3))
        dt = np.float64 if g % 3 else np.float32
        ta, tb = (111 if r.random() < 0.6 else 112), (111 if r.random() < 0.6 else 112)
        A = np.asfortranarray(r.uniform(-0.5, 0.5, (m, k) if ta == 111 else (k, m)).astype(dt))
        B = np.asfortranarray(r.uniform(-0.5, 0.5, (k, n) if tb == 111 else (n, k)).astype(dt))
        C = np.asfortranarray(r.uniform(-0.5, 0.5, (m, n)).astype(dt))
        dA, dB, dC = xk.DeviceArray(A), xk.DeviceArray(B), xk.DeviceArray(C)
        alpha, beta = 1.5, -0.5
        probs.append(xk.BatchProblem(xk.BATCH_GEMM, 0 if dt == np.float64 else 1, ta, tb, m, n, k, alpha, beta,
                                     dA.ptr, A.shape[0], dB.ptr, B.shape[0], dC.ptr, m))
        keep += [dA, dB, dC]
        checks.append(("gemm", dt, (ta, tb, alpha, beta, A, B, C), dC))
    for t in range(ntrsm):
        m, n = int(r.integers(1, maxtri)), int(r.integers(1, 130))
        dt = np.float64 if t % 3 else np.float32
        L = np.tril(r.uniform(-1, 1, (m, m)))
        np.fill_diagonal(L, 1.0 + np.abs(L).sum(axis=1))  # diagonally dominant
        L = np.asfortranarray(L.astype(dt))
        B = np.asfortranarray(r.uniform(-1, 1, (m, n)).astype(dt))
        dL, dB = xk.DeviceArray(L), xk.DeviceArray(B)
        probs.append(xk.BatchProblem(xk.BATCH_TRSM, 0 if dt == np.float64 else 1, 111, 111, m, n, m, 0.75, 0.0,
                                     dL.ptr, m, None, 0, dB.ptr, m))
        keep += [dL, dB]
        checks.append(("trsm", dt, (0.75, L, B), dB))
    arr = (xk.BatchProblem * len(probs))(*probs)
    return arr, keep, checks


def verify(orc, checks):
    for kind, dt, args, dout in checks:
        got = dout.download()
        if kind == "gemm":
            ta, tb, alpha, beta, A, B, C = args
            if dt == np.float64:
                ref = orc.dgemm_tiled(ta, tb, alpha, A, B, beta, C.copy(order="F"), 64)
                assert orc.gemm_residual(alpha, A, B, beta, C, ref, got) <= 3.0
            else:
                opA = A.astype(np.float64) if ta == 111 else A.astype(np.float64).T
                opB = B.astype(np.float64) if tb == 111 else B.astype(np.float64).T
                ref = alpha * opA @ opB + beta * C.astype(np.float64)
                assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
        else:
            alpha, L, B = args
            if dt == np.float64:
                ref = orc.dtrsm_tiled(141, 122, 111, 131, alpha, L, B.copy(order="F"), 64)
                assert orc.trsm_residual(L, 122, 131, ref, got) <= 10.0
            else:
                ref = np.linalg.solve(L.astype(np.float64), alpha * B.astype(np.float64))
                assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_batch_mixed_eager(xk, orc):
    r = rng(2024)
    arr, keep, checks = make_batch(xk, r, 60, 20)
    L = xk.lib()
    plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
    assert plan
    n_items = ctypes.c_int()
    L.xkblas_batch_info(plan, ctypes.byref(n_items), None, None, None)
    assert n_items.value > len(arr) // 2
    assert L.xkblas_batch_run(plan) == 0
    L.xkblas_sync()
    verify(orc, checks)
    L.xkblas_batch_destroy(plan)
    for d in keep:
        d.free()


def test_batch_graph_replays(xk, orc):
    """Capture the launch pair in a hipGraph; every replay (with the inputs
    re-uploaded eagerly in between) must reproduce the eager result."""
    r = rng(7)
    arr, keep, checks = make_batch(xk, r, 30, 10, maxdim=130, maxtri=200)
    L = xk.lib()
    plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
    outs = [c[3] for c in checks]
    inits = []
    for kind, dt, args, dout in checks:
        inits.append(args[6] if kind == "gemm" else args[2])
    L.xkblas_batch_run(plan)
    L.xkblas_sync()
    eager = [d.download() for d in outs]
    assert L.xkblas_batch_use_graph(plan, 1) == 0
    for rep in range(3):
        for d, h in zip(outs, inits):
            d.upload(h)
        L.xkblas_batch_run(plan)
        L.xkblas_sync()
        for d, e in zip(outs, eager):
            assert np.array_equal(d.download(), e), "replay %d differs from the eager run" % rep
    L.xkblas_batch_use_graph(plan, 0)
    L.xkblas_batch_destroy(plan)
    for d in keep:
        d.free()


def test_batch_edge_sizes(xk, orc):
    """Degenerate and tile-boundary shapes: 1, 63, 64, 65 and k = 0."""
    L = xk.lib()
    r = rng(5)
    probs, keep, checks = [], [], []
    for (m, n, k) in [(1, 1, 1), (64, 64, 64), (65, 63, 1), (63, 65, 129), (5, 7, 0)]:
        A = np.asfortranarray(r.uniform(-1, 1, (m, max(k, 1))))
        B = np.asfortranarray(r.uniform(-1, 1, (max(k, 1), n)))
        C = np.asfortranarray(r.uniform(-1, 1, (m, n)))
        dA, dB, dC = xk.DeviceArray(A), xk.DeviceArray(B), xk.DeviceArray(C)
        probs.append(xk.BatchProblem(0, 0, 111, 111, m, n, k, 2.0, 0.5, dA.ptr, m, dB.ptr, max(k, 1), dC.ptr, m))
        keep += [dA, dB, dC]
        ref = 2.0 * A[:, :k] @ B[:k, :] + 0.5 * C
        checks.append((dC, ref))
    arr = (xk.BatchProblem * len(probs))(*probs)
    plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
    L.xkblas_batch_run(plan)
    L.xkblas_sync()
    for d, ref in checks:
        np.testing.assert_allclose(d.download(), ref, rtol=0, atol=1e-13)
    L.xkblas_batch_destroy(plan)
    for d in keep:
        d.free()


def test_batch_trsm_edge_sizes(xk, orc):
    """TRSM block boundaries (m = 1, 16, 17, 63, 64, 65, 128, 129, 200) and
    panel boundaries (n = 1, 64, 65), FP64 and FP32, alpha in {-2, 0, 1}."""
    L = xk.lib()
    r = rng(6)
    probs, keep, checks = [], [], []
    cases = [(1, 1), (16, 64), (17, 65), (63, 3), (64, 64), (65, 1), (128, 65), (129, 64), (200, 130)]
    for t, (m, n) in enumerate(cases):
        for dt, code in ((np.float64, 0), (np.float32, 1)):
            alpha = (-2.0, 0.0, 1.0)[t % 3]
            Lm = np.tril(r.uniform(-1, 1, (m, m)))
            np.fill_diagonal(Lm, 1.0 + np.abs(Lm).sum(axis=1))
            Lm = np.asfortranarray(Lm.astype(dt))
            B = np.asfortranarray(r.uniform(-1, 1, (m, n)).astype(dt))
            dL, dB = xk.DeviceArray(Lm), xk.DeviceArray(B)
            probs.append(xk.BatchProblem(xk.BATCH_TRSM, code, 111, 111, m, n, m, alpha, 0.0, dL.ptr, m, None, 0,
                                         dB.ptr, m))
            keep += [dL, dB]
            ref = alpha * np.linalg.solve(Lm.astype(np.float64), B.astype(np.float64))
            checks.append((dB, ref, 1e-13 if dt == np.float64 else 2e-5))
    arr = (xk.BatchProblem * len(probs))(*probs)
    plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
    assert L.xkblas_batch_run(plan) == 0
    L.xkblas_sync()
    for d, ref, tol in checks:
        got = d.download().astype(np.float64)
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())
    L.xkblas_batch_destroy(plan)
    for d in keep:
        d.free()
