"""ZGEMM / CGEMM (reference blas/zgemm.c, gen z and c) through the C ABI:
every op(A) x op(B) combination of N, T, C against the textbook complex
product in NumPy, with the reference driver's normwise GEMM bar
(testing/ztesting_gemm.c:288-306, <= 3 in units of eps, complex moduli);
small Gaussian integers must be bit-exact; beta = 0 never reads C."""
import numpy as np
import pytest

from helpers import rng

pytestmark = pytest.mark.gpu

OPS = [111, 112, 113]
EPS = 2.0 ** -53


def _c(r, m, n, dtype=np.complex128):
    return np.asfortranarray((r.uniform(-0.5, 0.5, (m, n)) + 1j * r.uniform(-0.5, 0.5, (m, n))).astype(dtype))


def _op(X, t):
    return X if t == 111 else (X.T if t == 112 else X.conj().T)


def _inf(X):
    return float(np.abs(X).sum(axis=1).max()) if X.size else 0.0


def _bar(alpha, oA, oB, beta, C0, ref, out):
    K = oA.shape[1]
    norm = np.sqrt(K + 2.0) * abs(alpha) * _inf(oA) * _inf(oB) + np.sqrt(2.0) * abs(beta) * _inf(C0)
    return _inf(ref - out) / norm / EPS


@pytest.mark.parametrize("ta", OPS)
@pytest.mark.parametrize("tb", OPS)
@pytest.mark.parametrize("shape", [(130, 70, 90), (1, 1, 1), (257, 3, 300)])
def test_zgemm(xk, ta, tb, shape):
    M, N, K = shape
    r = rng(M + N + K + ta * 3 + tb)
    A = _c(r, M, K) if ta == 111 else _c(r, K, M)
    B = _c(r, K, N) if tb == 111 else _c(r, N, K)
    C0 = _c(r, M, N)
    alpha, beta = 1.5 - 0.5j, -0.5 + 0.25j
    C = C0.copy(order="F")
    xk.gemm(ta, tb, alpha, A, B, beta, C)
    xk.invalidate()
    ref = alpha * (_op(A, ta) @ _op(B, tb)) + beta * C0
    assert _bar(alpha, _op(A, ta), _op(B, tb), beta, C0, ref, C) <= 3.0


def test_zgemm_special_scalars_and_exact(xk):
    r = rng(61)
    M, N, K = 96, 80, 120
    A = np.asfortranarray(r.integers(-4, 5, (M, K)) + 1j * r.integers(-4, 5, (M, K)))
    B = np.asfortranarray(r.integers(-4, 5, (N, K)) + 1j * r.integers(-4, 5, (N, K)))
    C = np.full((M, N), np.nan + 1j * np.nan, order="F")
    xk.gemm(111, 113, 1.0, A, B, 0.0, C)  # beta = 0: C is not read
    xk.invalidate()
    assert np.array_equal(C, A @ B.conj().T)
    B = np.asfortranarray(B.T)
    C0 = np.asfortranarray(r.integers(-4, 5, (M, N)) + 1j * r.integers(-4, 5, (M, N)))
    C = C0.copy(order="F")
    nanA = np.full((M, K), np.nan, dtype=np.complex128, order="F")
    xk.gemm(111, 111, 0.0, nanA, B, 2.0 - 1.0j, C)  # alpha = 0: A, B are not read
    xk.invalidate()
    assert np.array_equal(C, (2.0 - 1.0j) * C0)


def test_zgemm_split_mode(xk):
    L = xk.lib()
    r = rng(67)
    M, N, K = 200, 150, 333
    A, B, C0 = _c(r, M, K), _c(r, N, K), _c(r, M, N)
    L.xkblas_set_modemath(2)
    try:
        C = C0.copy(order="F")
        xk.gemm(111, 113, 1.0 + 1.0j, A, B, 0.5, C)
    finally:
        L.xkblas_set_modemath(0)
        xk.invalidate()
    ref = (1.0 + 1.0j) * (A @ B.conj().T) + 0.5 * C0
    assert _bar(1.0 + 1.0j, A, B.conj().T, 0.5, C0, ref, C) <= 3.0


@pytest.mark.parametrize("ta,tb", [(111, 111), (113, 112)])
def test_cgemm(xk, ta, tb):
    r = rng(71 + ta + tb)
    M, N, K = 180, 90, 150
    A = _c(r, M, K, np.complex64) if ta == 111 else _c(r, K, M, np.complex64)
    B = _c(r, K, N, np.complex64) if tb == 111 else _c(r, N, K, np.complex64)
    C0 = _c(r, M, N, np.complex64)
    C = C0.copy(order="F")
    xk.gemm(ta, tb, np.complex64(0.5 + 2.0j), A, B, np.complex64(1.0), C)
    xk.invalidate()
    ref = (0.5 + 2.0j) * (_op(A.astype(np.complex128), ta) @ _op(B.astype(np.complex128), tb)) + C0
    assert np.abs(C - ref).max() / np.abs(ref).max() < 2e-6


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 113])
@pytest.mark.parametrize("NK", [(150, 70), (1, 5), (64, 200)])
def test_zherk(xk, uplo, trans, NK):
    N, K = NK
    r = rng(N + K + uplo + trans)
    A = _c(r, N, K) if trans == 111 else _c(r, K, N)
    C0 = _c(r, N, N)
    C = C0.copy(order="F")
    xk.herk(uplo, trans, 0.75, A, -1.5, C)
    xk.invalidate()
    oA = A if trans == 111 else A.conj().T
    ref = 0.75 * (oA @ oA.conj().T) - 1.5 * C0
    np.fill_diagonal(ref, ref.diagonal().real)
    mask = np.tril(np.ones((N, N), bool)) if uplo == 122 else np.triu(np.ones((N, N), bool))
    R = _inf(np.where(mask, ref - C, 0))
    assert R / ((_inf(oA) ** 2 + _inf(C0)) * max(N, 1) * EPS) <= 10.0
    assert np.all(C.diagonal().imag == 0.0)
    assert np.array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
def test_zhemm(xk, side, uplo):
    M, N = 140, 90
    ka = M if side == 141 else N
    r = rng(side + uplo)
    H = _c(r, ka, ka)
    H = H + H.conj().T
    np.fill_diagonal(H, H.diagonal().real)
    mask = np.tril(np.ones((ka, ka), bool)) if uplo == 122 else np.triu(np.ones((ka, ka), bool))
    A = np.asfortranarray(np.where(mask, H, np.nan))
    A[np.diag_indices(ka)] += 7j  # the imaginary part of the diagonal is not read
    B, C0 = _c(r, M, N), _c(r, M, N)
    alpha, beta = 0.5 + 1.0j, 1.0 - 0.5j
    C = C0.copy(order="F")
    xk.hemm(side, uplo, alpha, A, B, beta, C)
    xk.invalidate()
    ref = alpha * (H @ B if side == 141 else B @ H) + beta * C0
    oA, oB = (H, B) if side == 141 else (B, H)
    assert _bar(alpha, oA, oB, beta, C0, ref, C) <= 3.0


def test_cherk_chemm_fp32(xk):
    r = rng(79)
    N, K = 120, 60
    A = _c(r, N, K, np.complex64)
    C0 = _c(r, N, N, np.complex64)
    C = C0.copy(order="F")
    xk.herk(122, 111, 1.0, A, 1.0, C)
    xk.invalidate()
    A64 = A.astype(np.complex128)
    ref = A64 @ A64.conj().T + C0
    np.fill_diagonal(ref, ref.diagonal().real)
    assert np.abs(np.tril(C - ref)).max() / np.abs(ref).max() < 2e-6
    H = (C0 + C0.conj().T).astype(np.complex64)
    B = _c(r, N, 30, np.complex64)
    Y = np.zeros((N, 30), dtype=np.complex64, order="F")
    xk.hemm(141, 121, np.complex64(1.0), np.asfortranarray(H), B, np.complex64(0.0), Y)
    xk.invalidate()
    ref = H.astype(np.complex128) @ B.astype(np.complex128)
    assert np.abs(Y - ref).max() / np.abs(ref).max() < 2e-6


def _tri(n, uplo):
    return np.tril(np.ones((n, n), bool)) if uplo == 122 else np.triu(np.ones((n, n), bool))


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 113])
def test_zher2k(xk, uplo, trans):
    N, K = 130, 75
    r = rng(uplo + trans + 3)
    A = _c(r, N, K) if trans == 111 else _c(r, K, N)
    B = _c(r, N, K) if trans == 111 else _c(r, K, N)
    C0 = _c(r, N, N)
    alpha = 0.5 - 1.5j
    C = C0.copy(order="F")
    xk.her2k(uplo, trans, alpha, A, B, 2.0, C)
    xk.invalidate()
    oA, oB = (A, B) if trans == 111 else (A.conj().T, B.conj().T)
    ref = alpha * (oA @ oB.conj().T) + np.conj(alpha) * (oB @ oA.conj().T) + 2.0 * C0
    np.fill_diagonal(ref, ref.diagonal().real)
    m = _tri(N, uplo)
    assert _inf(np.where(m, ref - C, 0)) / ((abs(alpha) * _inf(oA) * _inf(oB) + _inf(C0)) * N * EPS) <= 10.0
    assert np.all(C.diagonal().imag == 0.0)
    assert np.array_equal(C[~m], C0[~m])


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
def test_zsyrk_zsyr2k(xk, uplo, trans):
    N, K = 120, 90
    r = rng(uplo + 5 * trans)
    A = _c(r, N, K) if trans == 111 else _c(r, K, N)
    B = _c(r, N, K) if trans == 111 else _c(r, K, N)
    C0 = _c(r, N, N)
    oA, oB = (A, B) if trans == 111 else (A.T, B.T)
    m = _tri(N, uplo)
    alpha, beta = 1.0 + 0.5j, -0.25 + 1.0j
    C = C0.copy(order="F")
    xk.syrk(uplo, trans, alpha, A, beta, C)
    xk.invalidate()
    ref = alpha * (oA @ oA.T) + beta * C0
    assert _inf(np.where(m, ref - C, 0)) / ((abs(alpha) * _inf(oA) ** 2 + abs(beta) * _inf(C0)) * N * EPS) <= 10.0
    assert np.array_equal(C[~m], C0[~m])
    C = C0.copy(order="F")
    xk.syr2k(uplo, trans, alpha, A, B, beta, C)
    xk.invalidate()
    ref = alpha * (oA @ oB.T + oB @ oA.T) + beta * C0
    assert _inf(np.where(m, ref - C, 0)) / ((abs(alpha) * _inf(oA) * _inf(oB) + abs(beta) * _inf(C0)) * N * EPS) <= 10.0
    assert np.array_equal(C[~m], C0[~m])


@pytest.mark.parametrize("uplo", [0, 121, 122])
def test_zgemmt(xk, uplo):
    N, K = 110, 140
    r = rng(uplo + 11)
    A, B, C0 = _c(r, N, K), _c(r, N, K), _c(r, N, N)
    C = C0.copy(order="F")
    xk.gemmt(uplo, 111, 113, 2.0j, A, B, 0.5, C)
    xk.invalidate()
    ref = 2.0j * (A @ B.conj().T) + 0.5 * C0
    m = np.ones((N, N), bool) if uplo == 0 else _tri(N, uplo)
    assert _bar(2.0j, A, B.conj().T, 0.5, C0, np.where(m, ref, 0), np.where(m, C, 0)) <= 3.0
    assert np.array_equal(C[~m], C0[~m])


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
def test_zsymm(xk, side, uplo):
    M, N = 100, 70
    ka = M if side == 141 else N
    r = rng(side + 2 * uplo)
    S = _c(r, ka, ka)
    S = S + S.T
    A = np.asfortranarray(np.where(_tri(ka, uplo), S, np.nan))
    B, C0 = _c(r, M, N), _c(r, M, N)
    C = C0.copy(order="F")
    xk.symm(side, uplo, 1.0 - 1.0j, A, B, 0.5j, C)
    xk.invalidate()
    ref = (1.0 - 1.0j) * (S @ B if side == 141 else B @ S) + 0.5j * C0
    oA, oB = (S, B) if side == 141 else (B, S)
    assert _bar(1.0 - 1.0j, oA, oB, 0.5j, C0, ref, C) <= 3.0


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", [111, 112, 113])
@pytest.mark.parametrize("diag", [131, 132])
def test_ztrmm(xk, side, uplo, ta, diag):
    M, N = 90, 60
    ka = M if side == 141 else N
    r = rng(side + uplo + ta + diag)
    T = np.where(_tri(ka, uplo), _c(r, ka, ka), 0)
    A = np.asfortranarray(np.where(_tri(ka, uplo), T, np.nan))
    if diag == 132:
        np.fill_diagonal(T, 1.0)
        np.fill_diagonal(A, np.nan)
    B0 = _c(r, M, N)
    B = B0.copy(order="F")
    xk.trmm(side, uplo, ta, diag, 0.5 + 0.5j, A, B)
    xk.invalidate()
    oT = _op(T, ta)
    ref = (0.5 + 0.5j) * (oT @ B0 if side == 141 else B0 @ oT)
    oA, oB = (oT, B0) if side == 141 else (B0, oT)
    assert _bar(0.5 + 0.5j, oA, oB, 0.0, B0, ref, B) <= 3.0


def test_csyrk_ctrmm_fp32(xk):
    r = rng(83)
    N, K = 100, 50
    A = _c(r, N, K, np.complex64)
    C0 = _c(r, N, N, np.complex64)
    C = C0.copy(order="F")
    xk.syrk(121, 111, np.complex64(1.0j), A, np.complex64(1.0), C)
    xk.invalidate()
    A64 = A.astype(np.complex128)
    ref = 1.0j * (A64 @ A64.T) + C0
    assert np.abs(np.triu(C - ref)).max() / np.abs(ref).max() < 2e-6
    T = np.asfortranarray(np.triu(_c(r, N, N, np.complex64)))
    X = C0.copy(order="F")
    xk.trmm(141, 121, 113, 131, np.complex64(1.0), T, X)
    xk.invalidate()
    ref = T.astype(np.complex128).conj().T @ C0.astype(np.complex128)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 2e-6


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", [111, 112, 113])
@pytest.mark.parametrize("diag", [131, 132])
def test_ztrsm(xk, side, uplo, ta, diag):
    """The reference TRSM test matrix made complex (testing/ztesting_trsm.c:
    129-134: entries / n, unit diagonal); bar <= 10 as there."""
    M, N = 300, 170
    na = M if side == 141 else N
    r = rng(side * 7 + uplo + ta + diag)
    T = (r.uniform(0.0, 1.0, (na, na)) + 1j * r.uniform(0.0, 1.0, (na, na))) / na
    np.fill_diagonal(T, 1.0 + 0.5j)
    T = np.where(_tri(na, uplo), T, 0)
    A = np.asfortranarray(np.where(_tri(na, uplo), T, np.nan))
    Tm = T.copy()
    if diag == 132:
        np.fill_diagonal(Tm, 1.0)
    B0 = _c(r, M, N)
    X = B0.copy(order="F")
    alpha = 0.8 - 0.2j
    xk.lib().xkblas_set_param(128, 0)
    xk.trsm(side, uplo, ta, diag, alpha, A, X)
    xk.invalidate()
    oT = _op(Tm, ta)
    Xref = np.linalg.solve(oT, alpha * B0) if side == 141 else np.linalg.solve(oT.T, (alpha * B0).T).T
    err = _inf(X - Xref) / ((_inf(oT) + _inf(Xref)) * max(M, N) * EPS)
    assert err <= 10.0, err


def test_ctrsm_fp32(xk):
    r = rng(89)
    n, nrhs = 200, 40
    T = np.tril((r.uniform(0, 1, (n, n)) + 1j * r.uniform(0, 1, (n, n))) / n)
    np.fill_diagonal(T, 1.0)
    T = np.asfortranarray(T.astype(np.complex64))
    B0 = _c(r, n, nrhs, np.complex64)
    X = B0.copy(order="F")
    xk.trsm(141, 122, 113, 131, np.complex64(1.0), T, X)
    xk.invalidate()
    ref = np.linalg.solve(T.astype(np.complex128).conj().T, B0.astype(np.complex128))
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-5


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("n_nb", [(64, 64), (300, 128), (777, 256), (130, 512)])
def test_zpotrf(xk, uplo, n_nb):
    n, nb = n_nb
    r = rng(n + uplo)
    G = _c(r, n, n)
    H = np.asfortranarray(G @ G.conj().T + n * np.eye(n))
    A = H.copy(order="F")
    xk.lib().xkblas_set_param(nb, 0)
    info = xk.potrf(uplo, A)
    xk.invalidate()
    assert info == 0
    F = np.tril(A) if uplo == 122 else np.triu(A)
    R = H - (F @ F.conj().T if uplo == 122 else F.conj().T @ F)
    assert _inf(R) / (n * _inf(H) * EPS) <= 10.0
    assert np.all(np.abs(F.diagonal().imag) == 0.0) and np.all(F.diagonal().real > 0)
    other = np.triu(np.ones((n, n), bool), 1) if uplo == 122 else np.tril(np.ones((n, n), bool), -1)
    assert np.array_equal(A[other], H[other])
    Lref = np.linalg.cholesky(H)
    Fref = Lref if uplo == 122 else Lref.conj().T
    assert np.abs(F - Fref).max() <= 1e-12 * np.abs(Fref).max()


@pytest.mark.parametrize("bad", [0, 150, 299])
def test_zpotrf_info(xk, bad):
    n = 300
    A = np.asfortranarray(np.diag(np.linspace(1.0, 2.0, n)).astype(np.complex128))
    A[bad, bad] = -1.0
    xk.lib().xkblas_set_param(128, 0)
    info = xk.potrf(122, A)
    xk.invalidate()
    assert info == bad + 1


def test_cpotrf_fp32(xk):
    r = rng(97)
    n = 200
    G = _c(r, n, n, np.complex64)
    H = (G @ G.conj().T + n * np.eye(n)).astype(np.complex64)
    A = np.asfortranarray(H.copy())
    xk.lib().xkblas_set_param(64, 0)
    assert xk.potrf(122, A) == 0
    xk.invalidate()
    F = np.tril(A).astype(np.complex128)
    assert np.abs(H - F @ F.conj().T).max() / np.abs(H).max() < 1e-5
