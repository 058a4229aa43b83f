"""Split modes (XKBLAS_SPLIT_BF16_MATH, XKBLAS_SPLIT_FP8_MATH): FP64-accurate
products from BF16 or FP8 residue GEMMs with exact FP32 accumulation and CRT
recombination.  Every test runs in both modes.

Bars:
  * integer inputs that fit the scaled width: the split product is exact, so
    it must equal the exact integer result bit for bit;
  * general inputs (2^u row/column scaling, u ~ U[-8, 8]): normwise relative
    error vs the double-double oracle <= 1e-14, and componentwise error
    within 2^-45 of (|alpha| |A| |B| + |beta| |C|);
  * SYRK / POTRF in split mode: the reference test bounds.
"""
import numpy as np
import pytest

from helpers import rng, uniform, small_int, spd

pytestmark = pytest.mark.gpu

SPLIT_BF16, SPLIT_FP8 = 2, 3


@pytest.fixture(params=[SPLIT_BF16, SPLIT_FP8], ids=["bf16", "fp8"])
def split_mode(xk, request):
    L = xk.lib()
    L.xkblas_set_modemath(request.param)
    L.mode = request.param
    yield L
    L.xkblas_set_modemath(0)
    L.xkblas_set_split_count(0)
    xk.invalidate()


def scaled(r, m, n, rows=True):
    X = r.uniform(-0.5, 0.5, size=(m, n))
    if rows:
        X *= np.exp2(r.uniform(-8, 8, size=m))[:, None]
    else:
        X *= np.exp2(r.uniform(-8, 8, size=n))[None, :]
    return np.asfortranarray(X)


@pytest.mark.parametrize("ta", [111, 112])
@pytest.mark.parametrize("tb", [111, 112])
def test_split_integer_exact(xk, split_mode, ta, tb):
    r = rng(100 + ta + tb)
    M, N, K = 300, 200, 333
    lim = 2**18
    A = small_int(r, M, K, -lim, lim) if ta == 111 else small_int(r, K, M, -lim, lim)
    B = small_int(r, K, N, -lim, lim) if tb == 111 else small_int(r, N, K, -lim, lim)
    C0 = small_int(r, M, N, -lim, lim)
    C = C0.copy(order="F")
    xk.gemm(ta, tb, 1.0, A, B, 2.0, C)
    opA = A if ta == 111 else A.T
    opB = B if tb == 111 else B.T
    exact = opA.astype(np.int64) @ opB.astype(np.int64) + 2 * C0.astype(np.int64)
    assert np.abs(exact).max() < 2**53
    assert np.array_equal(C, exact.astype(np.float64))


@pytest.mark.parametrize("shape", [(256, 256, 256), (513, 300, 1000), (100, 700, 64), (1024, 1024, 4096)])
def test_split_accuracy_vs_dd(xk, orc, split_mode, shape):
    M, N, K = shape
    r = rng(sum(shape))
    A = scaled(r, M, K, rows=True)
    B = scaled(r, K, N, rows=False)
    C0 = uniform(r, M, N)
    alpha, beta = 1.5, -0.5
    C = C0.copy(order="F")
    xk.gemm(111, 111, alpha, A, B, beta, C)
    hi, lo = orc.dgemm_dd(111, 111, alpha, A, B, beta, C0)
    err = np.abs(C - hi - lo)
    normwise = np.linalg.norm(C - hi - lo) / np.linalg.norm(hi)
    assert normwise <= 1e-14, normwise
    bound = np.abs(alpha) * np.abs(A) @ np.abs(B) + np.abs(beta) * np.abs(C0)
    assert np.all(err <= 2.0**-45 * bound)


@pytest.mark.parametrize("ta", [111, 112])
def test_split_extreme_row_scales(xk, orc, split_mode, ta):
    """Row/column shifts at the edges of the 2-norm rule: rows of op(A) at
    2^600 (squares overflow) and 2^-700 (squares underflow) take their shift
    from the row maximum instead; rows of equal-magnitude entries are the
    2-norm bound's worst case.  beta = 0 keeps every entry at its own scale
    (2^-1000 .. 2^900); componentwise error within 2^-44 of |A| |B|."""
    r = rng(404 + ta)
    M, N, K = 96, 64, 3000
    X = r.uniform(-0.5, 0.5, size=(M, K))
    X[:12] = np.where(X[:12] < 0, -0.25, 0.25)  # equal magnitudes
    X *= np.exp2(r.choice([-700.0, 0.0, 600.0], size=M))[:, None]
    B = r.uniform(-0.5, 0.5, size=(K, N))
    B[:, :6] = np.where(B[:, :6] < 0, -1.0, 1.0)
    B *= np.exp2(r.choice([-300.0, 0.0, 300.0], size=N))[None, :]
    A = np.asfortranarray(X if ta == 111 else X.T)
    B = np.asfortranarray(B)
    C = np.zeros((M, N), order="F")
    xk.gemm(ta, 111, 1.0, A, B, 0.0, C)
    hi, lo = orc.dgemm_dd(ta, 111, 1.0, A, B, 0.0, np.zeros((M, N), order="F"))
    bound = np.abs(X) @ np.abs(B)
    assert np.all(np.isfinite(C))
    assert np.all(np.abs(C - hi - lo) <= 2.0**-44 * bound)


def test_split_count_controls_accuracy(xk, orc, split_mode):
    r = rng(77)
    M = N = K = 512
    A, B = scaled(r, M, K), scaled(r, K, N, rows=False)
    Z = np.zeros((M, N), order="F")
    hi, lo = orc.dgemm_dd(111, 111, 1.0, A, B, 0.0, Z)
    errs = []
    for ns in ((9, 11, 13) if split_mode.mode == SPLIT_BF16 else (12, 15, 20)):
        split_mode.xkblas_set_split_count(ns)
        C = np.zeros((M, N), order="F")
        xk.gemm(111, 111, 1.0, A, B, 0.0, C)
        xk.invalidate()
        errs.append(np.linalg.norm(C - hi - lo) / np.linalg.norm(hi))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= 1e-15


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
def test_split_syrk(xk, orc, split_mode, uplo, trans):
    r = rng(5 + uplo + trans)
    N, K = 300, 257
    A = uniform(r, N, K) if trans == 111 else uniform(r, K, N)
    C0 = orc.plgsy(0.0, N, 51)
    Cref = orc.dsyrk_tiled(uplo, trans, 0.75, A, 0.25, C0.copy(order="F"), 128)
    C = C0.copy(order="F")
    xk.syrk(uplo, trans, 0.75, A, 0.25, C)
    assert orc.syrk_residual(A, C0, Cref, C, uplo) <= 10.0
    other = np.triu(np.ones((N, N)), 1) if uplo == 122 else np.tril(np.ones((N, N)), -1)
    assert np.array_equal(C[other > 0], C0[other > 0])


@pytest.mark.parametrize("uplo", [121, 122])
def test_split_potrf(xk, orc, split_mode, uplo):
    n, nb = 1000, 256
    A0 = spd(rng(n), n)
    Lref = A0.copy(order="F")
    assert orc.dpotrf_tiled(uplo, Lref, nb) == 0
    A = A0.copy(order="F")
    split_mode.xkblas_set_param(nb, 0)
    assert xk.potrf(uplo, A) == 0
    assert orc.potrf_residual(A0, A, uplo) <= 10.0
    tri = np.tril if uplo == 122 else np.triu
    np.testing.assert_allclose(tri(A), tri(Lref), rtol=0, atol=1e-12 * np.abs(Lref).max())


def test_split_trsm(xk, orc, split_mode):
    from helpers import tri_matrix
    r = rng(12)
    M, N = 600, 300
    A = tri_matrix(r, M, True, False)
    B0 = uniform(r, M, N, 0.0, 1.0)
    Xref = orc.dtrsm_tiled(141, 122, 111, 131, 0.8, A, B0.copy(order="F"), 128)
    X = B0.copy(order="F")
    split_mode.xkblas_set_param(128, 0)
    xk.trsm(141, 122, 111, 131, 0.8, A, X)
    assert orc.trsm_residual(A, 122, 131, Xref, X) <= 10.0


def test_split_long_inner_dimension(xk, orc, split_mode):
    """k beyond one FP8 launch's exact range (32768) is summed over chunks."""
    r = rng(31)
    M, N, K = 96, 80, 40000
    A = scaled(r, M, K, rows=True)
    B = scaled(r, K, N, rows=False)
    C0 = uniform(r, M, N)
    C = C0.copy(order="F")
    xk.gemm(111, 111, -1.25, A, B, 0.5, C)
    hi, lo = orc.dgemm_dd(111, 111, -1.25, A, B, 0.5, C0)
    assert np.linalg.norm(C - hi - lo) / np.linalg.norm(hi) <= 1e-14
    lim = 2**12
    Ai, Bi = small_int(r, M, K, -lim, lim), small_int(r, K, N, -lim, lim)
    Ci = np.zeros((M, N), order="F")
    xk.gemm(111, 111, 1.0, Ai, Bi, 0.0, Ci)
    assert np.array_equal(Ci, (Ai.astype(np.int64) @ Bi.astype(np.int64)).astype(np.float64))
