"""Pins of the CPU oracle (no GPU).

Each check is chosen so that a plausible slip in the restatement (wrong
transpose, dropped beta, wrong tile offset, wrong solve order) fails it:
  * tiled drivers vs the vendor host BLAS / LAPACK (MKL), the library the
    reference itself calls on the CPU, under the reference tests' bounds;
  * small-integer inputs: every product and partial sum is exact in FP64, so
    tiled oracle, untiled oracle and host BLAS agree bit for bit;
  * known factors: POTRF of L0 L0^T with an integer L0 and diagonal 8;
  * the double-double GEMM against exact rational arithmetic;
  * the zplgsy generator against a big-integer restatement of its LCG.
"""
import numpy as np
import pytest

from helpers import rng, uniform, small_int, spd, tri_matrix

TR = [111, 112]


@pytest.fixture(scope="module")
def blas(orc):
    p = orc.hostblas()
    if p is None:
        pytest.skip("no host BLAS in this image")
    return p


@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("tb", TR)
@pytest.mark.parametrize("nb", [16, 64, 1000])
def test_gemm_tiled_vs_hostblas(orc, blas, ta, tb, nb):
    M, N, K = 97, 130, 75
    r = rng(ta * 3 + tb + nb)
    A = uniform(r, M, K) if ta == 111 else uniform(r, K, M)
    B = uniform(r, K, N) if tb == 111 else uniform(r, N, K)
    C0 = uniform(r, M, N)
    C1 = orc.dgemm_tiled(ta, tb, 1.5, A, B, -0.5, C0.copy(order="F"), nb)
    C2 = orc.host_dgemm(ta, tb, 1.5, A, B, -0.5, C0.copy(order="F"))
    assert orc.gemm_residual(1.5, A, B, -0.5, C0, C2, C1) <= 3.0


@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("tb", TR)
def test_gemm_small_int_bit_exact(orc, blas, ta, tb):
    M, N, K = 70, 41, 90
    r = rng(5)
    A = small_int(r, M, K) if ta == 111 else small_int(r, K, M)
    B = small_int(r, K, N) if tb == 111 else small_int(r, N, K)
    C0 = small_int(r, M, N)
    C1 = orc.dgemm_tiled(ta, tb, 2.0, A, B, -3.0, C0.copy(order="F"), 32)
    C2 = orc.host_dgemm(ta, tb, 2.0, A, B, -3.0, C0.copy(order="F"))
    opA = A if ta == 111 else A.T
    opB = B if tb == 111 else B.T
    exact = 2 * (opA.astype(np.int64) @ opB.astype(np.int64)) - 3 * C0.astype(np.int64)
    assert np.array_equal(C1, exact) and np.array_equal(C2, exact)


def test_gemm_beta_applied_once(orc):
    """beta must multiply C once (k == 0 tile), not once per k tile."""
    r = rng(6)
    A, B, C0 = small_int(r, 8, 40), small_int(r, 40, 8), small_int(r, 8, 8)
    C = orc.dgemm_tiled(111, 111, 1.0, A, B, 2.0, C0.copy(order="F"), 8)  # 5 k tiles
    assert np.array_equal(C, A @ B + 2 * C0)


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", TR)
def test_syrk_tiled_vs_hostblas(orc, blas, uplo, trans):
    N, K = 101, 67
    r = rng(uplo + trans)
    A = uniform(r, N, K) if trans == 111 else uniform(r, K, N)
    C0 = orc.plgsy(0.0, N, 51)
    C1 = orc.dsyrk_tiled(uplo, trans, 0.5, A, 2.0, C0.copy(order="F"), 32)
    C2 = orc.host_dsyrk(uplo, trans, 0.5, A, 2.0, C0.copy(order="F"))
    assert orc.syrk_residual(A, C0, C2, C1, uplo) <= 10.0
    other = np.triu(np.ones((N, N)), 1) if uplo == 122 else np.tril(np.ones((N, N)), -1)
    assert np.array_equal(C1[other > 0], C0[other > 0])


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("diag", [131, 132])
def test_trsm_tiled_vs_hostblas(orc, blas, side, uplo, ta, diag):
    M, N = 90, 70
    na = M if side == 141 else N
    r = rng(side * 7 + uplo * 5 + ta * 3 + diag)
    A = tri_matrix(r, na, uplo == 122, diag == 132)
    B0 = uniform(r, M, N, 0.0, 1.0)
    X1 = orc.dtrsm_tiled(side, uplo, ta, diag, 0.7, A, B0.copy(order="F"), 32)
    X2 = orc.host_dtrsm(side, uplo, ta, diag, 0.7, A, B0.copy(order="F"))
    assert orc.trsm_residual(A, uplo, diag, X2, X1) <= 10.0


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", TR)
def test_trsm_tiled_solves_exactly(orc, side, uplo, ta):
    """Unit triangular +-1 bidiagonal T and integer X: B = alpha^-1 op(T) X
    (or X op(T)) is exact, and the tiled solve must return X bit for bit."""
    n, nrhs = 40, 24
    r = rng(side + uplo + ta)
    T = np.zeros((n, n))
    np.fill_diagonal(T, 1.0)
    off = r.integers(-1, 2, size=n - 1).astype(np.float64)
    if uplo == 122:
        T[np.arange(1, n), np.arange(n - 1)] = off
    else:
        T[np.arange(n - 1), np.arange(1, n)] = off
    T = np.asfortranarray(T)
    opT = T if ta == 111 else T.T
    if side == 141:
        X = small_int(r, n, nrhs)
        B = np.asfortranarray(opT @ X / 2.0)
    else:
        X = small_int(r, nrhs, n)
        B = np.asfortranarray(X @ opT / 2.0)
    orc.dtrsm_tiled(side, uplo, ta, 132, 2.0, T, B, 16)
    assert np.array_equal(B, X)


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("nb", [16, 50, 200])
def test_potrf_tiled_vs_lapack(orc, blas, uplo, nb):
    n = 150
    A0 = spd(rng(nb), n)
    L1 = A0.copy(order="F")
    L2 = A0.copy(order="F")
    assert orc.dpotrf_tiled(uplo, L1, nb) == 0
    assert orc.host_dpotrf(uplo, L2) == 0
    tri = np.tril if uplo == 122 else np.triu
    np.testing.assert_allclose(tri(L1), tri(L2), rtol=0, atol=1e-13 * np.abs(L2).max())
    assert orc.potrf_residual(A0, L1, uplo) <= 10.0


@pytest.mark.parametrize("uplo", [121, 122])
def test_potrf_tiled_known_factor(orc, uplo):
    n = 64
    r = rng(2)
    L0 = np.tril(small_int(r, n, n, -2, 2), -1)
    np.fill_diagonal(L0, 8.0)
    A = np.asfortranarray(L0 @ L0.T)
    F = A.copy(order="F")
    assert orc.dpotrf_tiled(uplo, F, 16) == 0
    got = np.tril(F) if uplo == 122 else np.triu(F).T
    np.testing.assert_allclose(got, L0, rtol=0, atol=1e-12)


def test_potrf_tiled_info(orc, blas):
    n = 60
    A = spd(rng(3), n)
    A[37, 37] = -5.0
    B = A.copy(order="F")
    assert orc.dpotrf_tiled(122, A, 16) == 38
    assert orc.host_dpotrf(122, B) == 38


def test_dd_gemm_exact(orc):
    """Entries with 30-bit integer significands: exact products need ~60 bits,
    sums ~66.  The double-double result must equal exact integer arithmetic
    to within one unit of its ~106-bit precision."""
    r = rng(8)
    M, N, K = 5, 4, 64
    A = np.asfortranarray(r.integers(-2**30, 2**30, size=(M, K)).astype(np.float64))
    B = np.asfortranarray(r.integers(-2**30, 2**30, size=(K, N)).astype(np.float64))
    C = np.zeros((M, N), order="F")
    hi, lo = orc.dgemm_dd(111, 111, 1.0, A, B, 0.0, C)
    Ai = [[int(A[i, k]) for k in range(K)] for i in range(M)]
    Bi = [[int(B[k, j]) for j in range(N)] for k in range(K)]
    for i in range(M):
        for j in range(N):
            exact = sum(Ai[i][k] * Bi[k][j] for k in range(K))
            got = int(hi[i, j]) + int(lo[i, j])
            assert abs(got - exact) <= abs(exact) * 2.0**-100 + 1
    # and FP64 alone cannot represent it
    assert np.any(lo != 0.0)


def test_plgsy_matches_lcg(orc):
    """testing/zplgsy.c restated with Python integers (mod 2^64) and float32."""
    N, seed, bump = 9, 51, 3.0
    A = orc.plgsy(bump, N, seed)
    a_mul, c_add, mask = 6364136223846793005, 1, (1 << 64) - 1

    def jump(n, s):
        a, c, ran = a_mul, c_add, s
        while n:
            if n & 1:
                ran = (a * ran + c) & mask
            c = (c * (a + 1)) & mask
            a = (a * a) & mask
            n >>= 1
        return ran

    E = np.zeros((N, N))
    for j in range(N):
        ran = jump(N, seed)
        for i in range(j, N):
            E[i, j] = np.float32(0.5) - np.float32(ran) * np.float32(5.4210108624275222e-20)
            ran = (a_mul * ran + c_add) & mask
    for j in range(N):
        E[j, j] += bump
        for i in range(j):
            E[i, j] = E[j, i]
    assert np.array_equal(A, E)


def test_flops_model(orc):
    L = orc.lib()
    assert L.oracle_flops_gemm(2048, 2048, 2048) == 2 * 2048**3
    assert L.oracle_flops_syrk(16384, 16384) == 16384 * 16384 * 16385
    n = 16384.0
    assert abs(L.oracle_flops_potrf(n) - (n**3 / 3 + n**2 / 2 + n / 6)) < 1e-6 * n**3
    assert L.oracle_flops_trsm(141, 100, 7) == 100 * 100 * 7
    assert L.oracle_flops_trsm_ref(141, 100, 7) == 7 * 100 * 101
