"""Allocation-free batched run (xkblas_batch_run_noalloc) through the C ABI,
against the plan path (xkblas_batch_create + xkblas_batch_run) on the same
problems.  Both build the same work list and run the same persistent kernel,
so every result must agree bit for bit.  Two allocation-free runs in a row,
the second on a smaller batch: the fixed device array still holds the first
run's lists, and the second run must neither read them nor touch the
problems it was not given."""
import ctypes

import numpy as np
import pytest

from helpers import rng

pytestmark = pytest.mark.gpu


def build(xk, r, ngemm, ntrsm):
    probs, outs = [], []
    for g in range(ngemm):
        m, n, k = (int(x) for x in r.integers(1, 300, 3))
        dt = np.float64 if g % 3 else np.float32
        ta, tb = (111 if r.random() < 0.6 else 112), (111 if r.random() < 0.6 else 112)
        A = r.uniform(-0.5, 0.5, (m, k) if ta == 111 else (k, m)).astype(dt)
        B = r.uniform(-0.5, 0.5, (k, n) if tb == 111 else (n, k)).astype(dt)
        C = np.asfortranarray(r.uniform(-0.5, 0.5, (m, n)).astype(dt))
        dA, dB, dC = xk.DeviceArray(A), xk.DeviceArray(B), xk.DeviceArray(C)
        probs.append(xk.BatchProblem(xk.BATCH_GEMM, 0 if dt == np.float64 else 1, ta, tb, m, n, k, 1.5, -0.5,
                                     dA.ptr, A.shape[0], dB.ptr, B.shape[0], dC.ptr, m))
        outs.append((dC, C, [dA, dB]))
    for t in range(ntrsm):
        m, n = int(r.integers(1, 700)), int(r.integers(1, 150))
        dt = np.float64 if t % 3 else np.float32
        L = np.tril(r.uniform(-1, 1, (m, m)))
        np.fill_diagonal(L, 1.0 + np.abs(L).sum(axis=1))
        L = L.astype(dt)
        B = np.asfortranarray(r.uniform(-1, 1, (m, n)).astype(dt))
        dL, dB = xk.DeviceArray(L), xk.DeviceArray(B)
        probs.append(xk.BatchProblem(xk.BATCH_TRSM, 0 if dt == np.float64 else 1, 111, 111, m, n, m, 0.75, 0.0,
                                     dL.ptr, m, None, 0, dB.ptr, m))
        outs.append((dB, B, [dL]))
    return probs, outs


def reset(L, outs):
    L.xkblas_sync()
    for d, h0, _ in outs:
        d.upload(h0)


def results(L, outs):
    L.xkblas_sync()
    return [d.download() for d, _, _ in outs]


def test_noalloc_matches_plan_twice(xk):
    L = xk.lib()
    r = rng(7001)
    probs, outs = build(xk, r, 48, 16)
    # interleave the kinds so a prefix holds both
    order = [i for pair in zip(range(16), range(48, 64)) for i in pair] + list(range(16, 48))
    probs = [probs[i] for i in order]
    outs = [outs[i] for i in order]
    arr = (xk.BatchProblem * len(probs))(*probs)
    try:
        plan = L.xkblas_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(arr))
        assert plan
        assert L.xkblas_batch_run(plan) == 0
        ref = results(L, outs)
        L.xkblas_batch_destroy(plan)

        # first allocation-free run: the whole batch
        reset(L, outs)
        assert L.xkblas_batch_run_noalloc(ctypes.cast(arr, ctypes.c_void_p), len(arr)) == 0
        got = results(L, outs)
        for g, w in zip(got, ref):
            assert np.array_equal(g, w)

        # second run: a prefix of 20 problems, inputs reset; the others keep
        # the first run's results
        small = 20
        reset(L, outs[:small])
        sub = (xk.BatchProblem * small)(*probs[:small])
        assert L.xkblas_batch_run_noalloc(ctypes.cast(sub, ctypes.c_void_p), small) == 0
        got2 = results(L, outs)
        for g, w in zip(got2, ref):
            assert np.array_equal(g, w)
        assert any(p.kind == xk.BATCH_TRSM for p in probs[:small])
        assert any(p.kind == xk.BATCH_GEMM for p in probs[:small])
    finally:
        L.xkblas_sync()
        for d, _, keep in outs:
            d.free()
            for k in keep:
                k.free()


def test_noalloc_rejects_oversize(xk):
    L = xk.lib()
    big = (xk.BatchProblem * 8193)()
    assert L.xkblas_batch_run_noalloc(ctypes.cast(big, ctypes.c_void_p), 8193) == -2
    # 8192 GEMMs with C 2048 x 2048: within the problem limit, but 1024 tiles
    # each is beyond the item limit; counted on the host, nothing runs
    many = (xk.BatchProblem * 8192)()
    for p in many:
        p.kind, p.dtype, p.m, p.n, p.k = xk.BATCH_GEMM, 0, 2048, 2048, 1
    assert L.xkblas_batch_run_noalloc(ctypes.cast(many, ctypes.c_void_p), 8192) == -3
    assert L.xkblas_batch_run_noalloc(None, 3) == -1
