"""GEMMT, SYR2K, SYMM and TRMM through the C ABI (reference blas/zgemmt.c,
zsyr2k.c, zsymm.c, ztrmm.c; BLAS semantics of dgemmt/dsyr2k/dsymm/dtrmm).

Each expected value is the textbook definition evaluated by NumPy in FP64;
the bars are the reference test drivers' normwise ones (GEMM <= 3 eps with
testing/ztesting_gemm.c:288-306's norm, SYR2K like SYRK <= 10,
testing/ztesting_syrk.c:225-235).  Small-integer inputs make every product
exact, so those cases must match bit for bit.  Entries a routine must not
read hold NaN; entries it must not write are compared unchanged.  Tests that
make several calls drop the device mirrors after each (the tile cache is keyed
by host pointer, and a new array may take a freed one's address).
"""
import numpy as np
import pytest

from helpers import rng, uniform, small_int

pytestmark = pytest.mark.gpu

NT, TT = 111, 112
UP, LO = 121, 122
LEFT, RIGHT = 141, 142
NONUNIT, UNIT = 131, 132
EPS = 2.0 ** -53


def _inf(X):
    return float(np.abs(X).sum(axis=1).max()) if X.size else 0.0


def _op(X, t):
    return X if t == NT else X.T


def _tri_mask(n, uplo):
    return np.tril(np.ones((n, n), bool)) if uplo == LO else np.triu(np.ones((n, n), bool))


def _gemm_bar(alpha, opA, opB, beta, C0, ref, out):
    K = opA.shape[1]
    norm = np.sqrt(K + 2.0) * abs(alpha) * _inf(opA) * _inf(opB) + np.sqrt(2.0) * abs(beta) * _inf(C0)
    return _inf(ref - out) / norm / EPS


@pytest.mark.parametrize("uplo", [0, UP, LO])
@pytest.mark.parametrize("ta", [NT, TT])
@pytest.mark.parametrize("tb", [NT, TT])
@pytest.mark.parametrize("NK", [(300, 129), (1, 1), (130, 500)])
def test_dgemmt(xk, uplo, ta, tb, NK):
    N, K = NK
    r = rng(N + K + ta + 3 * tb + uplo)
    A = uniform(r, N, K) if ta == NT else uniform(r, K, N)
    B = uniform(r, K, N) if tb == NT else uniform(r, N, K)
    C0 = uniform(r, N, N)
    C = C0.copy(order="F")
    xk.gemmt(uplo, ta, tb, 1.5, A, B, -0.5, C)
    xk.invalidate()
    ref = 1.5 * _op(A, ta) @ _op(B, tb) - 0.5 * C0
    mask = np.ones((N, N), bool) if uplo == 0 else _tri_mask(N, uplo)
    assert _gemm_bar(1.5, _op(A, ta), _op(B, tb), -0.5, C0, np.where(mask, ref, 0), np.where(mask, C, 0)) <= 3.0
    assert np.array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("trans", [NT, TT])
@pytest.mark.parametrize("NK", [(257, 100), (64, 300), (3, 1)])
def test_dsyr2k(xk, orc, uplo, trans, NK):
    N, K = NK
    r = rng(7 * N + K + uplo + trans)
    A = uniform(r, N, K) if trans == NT else uniform(r, K, N)
    B = uniform(r, N, K) if trans == NT else uniform(r, K, N)
    C0 = uniform(r, N, N)
    C = C0.copy(order="F")
    xk.syr2k(uplo, trans, 0.75, A, B, 1.25, C)
    xk.invalidate()
    oA, oB = (A, B) if trans == NT else (A.T, B.T)
    ref = 0.75 * (oA @ oB.T + oB @ oA.T) + 1.25 * C0
    mask = _tri_mask(N, uplo)
    R = _inf(np.where(mask, ref - C, 0))
    assert R / ((_inf(oA) * _inf(oB) + _inf(C0)) * N * EPS) <= 10.0
    assert np.array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("MN", [(300, 170), (1, 7), (65, 1)])
def test_dsymm(xk, side, uplo, MN):
    M, N = MN
    ka = M if side == LEFT else N
    r = rng(M * 3 + N + side + uplo)
    S = uniform(r, ka, ka)
    S = np.asfortranarray(S + S.T)
    A = S.copy(order="F")
    A[~_tri_mask(ka, uplo)] = np.nan  # the other triangle is not read
    B, C0 = uniform(r, M, N), uniform(r, M, N)
    C = C0.copy(order="F")
    xk.symm(side, uplo, 2.0, A, B, 0.5, C)
    xk.invalidate()
    ref = 2.0 * (S @ B if side == LEFT else B @ S) + 0.5 * C0
    opA, opB = (S, B) if side == LEFT else (B, S)
    assert _gemm_bar(2.0, opA, opB, 0.5, C0, ref, C) <= 3.0


@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("ta", [NT, TT])
@pytest.mark.parametrize("diag", [NONUNIT, UNIT])
def test_dtrmm(xk, side, uplo, ta, diag):
    M, N = 300, 170
    ka = M if side == LEFT else N
    r = rng(side + uplo + ta + diag)
    T = uniform(r, ka, ka)
    mask = _tri_mask(ka, uplo)
    T[~mask] = 0.0
    A = T.copy(order="F")
    A[~mask] = np.nan
    if diag == UNIT:
        np.fill_diagonal(T, 1.0)
        np.fill_diagonal(A, np.nan)  # a unit diagonal is not read
    B0 = uniform(r, M, N)
    B = B0.copy(order="F")
    xk.trmm(side, uplo, ta, diag, -1.5, A, B)
    xk.invalidate()
    ref = -1.5 * (_op(T, ta) @ B0 if side == LEFT else B0 @ _op(T, ta))
    opA, opB = (_op(T, ta), B0) if side == LEFT else (B0, _op(T, ta))
    assert _gemm_bar(-1.5, opA, opB, 0.0, B0, ref, B) <= 3.0


def test_small_integer_cases_are_exact(xk):
    r = rng(101)
    n, k = 200, 150
    A, B, C0 = small_int(r, n, k), small_int(r, n, k), small_int(r, n, n)
    C = C0.copy(order="F")
    xk.syr2k(LO, NT, 1.0, A, B, 2.0, C)
    xk.invalidate()
    ref = A @ B.T + B @ A.T + 2.0 * C0
    assert np.array_equal(np.tril(C), np.tril(ref))
    C = C0.copy(order="F")
    xk.gemmt(UP, NT, TT, 1.0, A, B, -1.0, C)
    xk.invalidate()
    assert np.array_equal(np.triu(C), np.triu(A @ B.T - C0))
    S = small_int(r, n, n)
    S = np.asfortranarray(np.tril(S) + np.tril(S, -1).T)
    Bm = small_int(r, n, 90)
    C = np.zeros((n, 90), order="F")
    xk.symm(LEFT, LO, 1.0, np.asfortranarray(np.tril(S)), Bm, 0.0, C)
    xk.invalidate()
    assert np.array_equal(C, S @ Bm)
    T = np.asfortranarray(np.triu(small_int(r, n, n)))
    X = Bm.copy(order="F")
    xk.trmm(LEFT, UP, TT, NONUNIT, 1.0, T, X)
    xk.invalidate()
    assert np.array_equal(X, T.T @ Bm)
    xk.invalidate()


def test_alpha_zero_reads_no_operand(xk):
    """alpha = 0: A and B are not referenced (they hold NaN)."""
    r = rng(103)
    n, k = 100, 40
    nanA = np.full((n, k), np.nan, order="F")
    C0 = uniform(r, n, n)
    C = C0.copy(order="F")
    xk.syr2k(UP, NT, 0.0, nanA, nanA, 3.0, C)
    xk.invalidate()
    assert np.array_equal(np.triu(C), np.triu(3.0 * C0))
    C = C0.copy(order="F")
    xk.gemmt(LO, NT, TT, 0.0, nanA, nanA, 0.0, C)
    xk.invalidate()
    assert np.array_equal(np.tril(C), np.zeros((n, n))) and np.array_equal(np.triu(C, 1), np.triu(C0, 1))
    X = uniform(r, n, 30)
    xk.trmm(LEFT, LO, NT, UNIT, 0.0, np.full((n, n), np.nan, order="F"), X)
    xk.invalidate()
    assert np.array_equal(X, np.zeros((n, 30)))
    xk.invalidate()


def test_split_mode_stays_fp64_accurate(xk):
    """Split-BF16 mode runs the products of GEMMT/SYR2K/SYMM/TRMM as CRT
    residue GEMMs: still within the FP64 bars."""
    L = xk.lib()
    r = rng(107)
    n, k = 384, 256
    A, B, C0 = uniform(r, n, k), uniform(r, n, k), uniform(r, n, n)
    S = uniform(r, n, n)
    S = np.asfortranarray(S + S.T)
    T = np.asfortranarray(np.tril(uniform(r, n, n)))
    L.xkblas_set_modemath(2)
    try:
        C = C0.copy(order="F")
        xk.gemmt(LO, NT, TT, 1.0, A, B, 0.5, C)
        xk.invalidate()
        assert _gemm_bar(1.0, A, B.T, 0.5, C0, np.tril(A @ B.T + 0.5 * C0), np.tril(C)) <= 3.0
        C = C0.copy(order="F")
        xk.syr2k(UP, NT, 1.0, A, B, 0.0, C)
        xk.invalidate()
        ref = A @ B.T + B @ A.T
        assert _inf(np.triu(ref - C)) / (_inf(A) * _inf(B) * n * EPS) <= 10.0
        Cs = np.zeros((n, n), order="F")
        xk.symm(RIGHT, UP, 1.0, S, C0, 0.0, Cs)
        xk.invalidate()
        assert _gemm_bar(1.0, C0, S, 0.0, C0, C0 @ S, Cs) <= 3.0
        X = C0.copy(order="F")
        xk.trmm(RIGHT, LO, NT, NONUNIT, 1.0, T, X)
        xk.invalidate()
        assert _gemm_bar(1.0, C0, T, 0.0, C0, C0 @ T, X) <= 3.0
    finally:
        L.xkblas_set_modemath(0)
        xk.invalidate()


def test_sgemmt_and_strmm_fp32(xk):
    r = rng(109)
    n, k = 257, 190
    A, B, C0 = (uniform(r, n, k, dtype=np.float32), uniform(r, n, k, dtype=np.float32),
                uniform(r, n, n, dtype=np.float32))
    C = C0.copy(order="F")
    xk.gemmt(LO, NT, TT, 1.0, A, B, 1.0, C)
    xk.invalidate()
    ref = A.astype(np.float64) @ B.T.astype(np.float64) + C0
    assert np.abs(np.tril(C - ref)).max() / np.abs(ref).max() < 1e-5
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))
    T = np.asfortranarray(np.triu(uniform(r, n, n, dtype=np.float32)))
    X = C0.copy(order="F")
    xk.trmm(LEFT, UP, NT, NONUNIT, 1.0, T, X)
    xk.invalidate()
    ref = T.astype(np.float64) @ C0.astype(np.float64)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-5
    xk.invalidate()
