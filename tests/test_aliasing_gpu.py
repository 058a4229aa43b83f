"""Operands that share host memory, and the POTRF info words.

BLAS lets read-only operands alias each other with different shapes (dgemm
with B pointing into A).  The tile cache keys mirrors by host pointer (as the
reference's descriptor map does, blas/xkblas.c:426-455), so the second
operand meets the first one's mirror under the same key with other
dimensions: it must get a mirror of its own while the first one stays alive
for the call, and device-only tiles of the old mirror must reach host memory
before the new one reads it.  In the split modes a product whose operands
are the same matrix read transposed is a SYRK; for real types ConjTrans is
Trans, so (Trans, ConjTrans) with B == A is A^T A^T, not A^T A.  Every
factored matrix owns an info word until it is invalidated; more matrices
than the initial 1024 words must each keep their own value.
"""
import ctypes

import numpy as np
import pytest

from helpers import rng, uniform

pytestmark = pytest.mark.gpu

N_, T_, C_ = 111, 112, 113


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dgemm_raw(L, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    al, be = np.array([alpha]), np.array([beta])
    return L.xkblas_dgemm_async(ta, tb, M, N, K, _p(al), A, lda, B, ldb, _p(be), _p(C), ldc)


@pytest.mark.parametrize("mode", [0, 2])
def test_dgemm_b_aliases_a_with_other_shape(xk, mode):
    """A = X[:300, :200], B = X[:200, :150] (same pointer, ld 400), twice in
    a row, the first time with X's mirror holding device-only tiles from a
    previous product into X[:300, :200]."""
    L = xk.lib()
    r = rng(11)
    LD, M, K, N = 400, 300, 200, 150
    X = uniform(r, LD, 300)
    P, Q = uniform(r, M, 64), uniform(r, 64, K)
    C0 = uniform(r, M, N)
    L.xkblas_set_param(128, 0)
    L.xkblas_set_modemath(mode)
    try:
        # X[:300, :200] <- P Q on the device only (no coherent yet)
        al, be = np.array([1.0]), np.array([0.0])
        assert L.xkblas_dgemm_async(N_, N_, M, K, 64, _p(al), _p(P), M, _p(Q), 64, _p(be), _p(X), LD) == 0
        Xh = X.copy(order="F")
        Xh[:M, :K] = P @ Q
        outs = []
        for _ in range(2):
            C = C0.copy(order="F")
            assert _dgemm_raw(L, N_, N_, M, N, K, 1.5, _p(X), LD, _p(X), LD, -0.5, C, M) == 0
            xk.coherent(C)
            outs.append(C)
        xk.coherent(X)
    finally:
        L.xkblas_set_modemath(0)
        xk.invalidate()
    ref = 1.5 * Xh[:M, :K] @ Xh[:K, :N] - 0.5 * C0
    tol = 1e-13 * np.abs(ref).max()
    np.testing.assert_allclose(X, Xh, rtol=0, atol=1e-14)
    for C in outs:
        np.testing.assert_allclose(C, ref, rtol=0, atol=tol)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("mode", [0, 2, 3])
@pytest.mark.parametrize("ta,tb", [(T_, C_), (C_, T_), (N_, C_), (C_, N_)])
def test_split_conjtrans_alias(xk, mode, ta, tb):
    """B == A (square), real type: op(A) op(A) with ConjTrans read as Trans."""
    L = xk.lib()
    r = rng(13 + ta + tb)
    n = 260
    A = uniform(r, n, n)
    C0 = uniform(r, n, n)
    L.xkblas_set_param(128, 0)
    L.xkblas_set_modemath(mode)
    try:
        C = C0.copy(order="F")
        assert _dgemm_raw(L, ta, tb, n, n, n, 1.0, _p(A), n, _p(A), n, 2.0, C, n) == 0
        xk.coherent(C)
    finally:
        L.xkblas_set_modemath(0)
        xk.invalidate()
    opA = A if ta == N_ else A.T
    opB = A if tb == N_ else A.T
    ref = opA @ opB + 2.0 * C0
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err <= 1e-14, err


def test_potrf_info_beyond_initial_slots(xk):
    """1100 matrices factored without invalidation: every third one has a
    negative pivot at order (i % 7) + 1; each keeps its own info."""
    L = xk.lib()
    r = rng(17)
    n = 16
    mats, want = [], []
    L.xkblas_set_param(64, 0)
    try:
        for i in range(1100):
            A = np.asfortranarray(np.diag(r.uniform(1.0, 2.0, n)))
            bad = i % 3 == 0
            if bad:
                A[i % 7, i % 7] = -1.0
            assert L.xkblas_dpotrf_async(122, n, _p(A), n) == 0
            mats.append(A)
            want.append((i % 7) + 1 if bad else 0)
        L.xkblas_sync()
        got = [L.xkblas_potrf_info(_p(A)) for A in mats]
    finally:
        xk.invalidate()
    assert got == want
