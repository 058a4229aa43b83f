"""CPU restatements of the complex diagonal-block kernels (xkblas_amd/csrc/
complex.hip), element by element as the kernels run them:

* cplx_trtri_leaf_kernel: lane j forward-substitutes column j of inv(L) in the
  lower view (Upper: L = U^T, inv(U) = inv(L)^T), unit or not;
* cplx_realform_kernel / cplx_fromreal_kernel: the complex POTRF's diagonal
  block packed to its 2n x 2n real form (lower triangle, from the stored
  triangle: Upper conjugated, the diagonal's imaginary parts dropped), the
  real Cholesky factor and its inverse read back into the stored triangle
  (Upper: U = L^H) and into inv(L) / inv(U); a failing real pivot 2q or
  2q + 1 reports complex order q + 1.

Each must give the inverse / the Hermitian factor of the textbook definition
(a conjugate missing in the view or the read-back, a swapped sign of the
imaginary block, or a wrong index, fails).
"""
import numpy as np
import pytest


def trtri_leaf(A, uplo, unit):
    n = A.shape[0]
    L = (lambda i, k: A[i, k]) if uplo == "L" else (lambda i, k: A[k, i])
    X = np.zeros((n, n), complex)
    for i in range(n):
        for j in range(n):  # lane j
            s = sum(L(i, k) * X[k, j] for k in range(i))
            if i < j:
                v = 0
            else:
                num = 1 if i == j else -s
                v = num if unit else num / L(i, i)
            X[i, j] = v
    return X if uplo == "L" else X.T


def realform(A, uplo):
    """cplx_realform_kernel: R[r][c] for r >= c (upper part left zero)."""
    n = A.shape[0]
    R = np.zeros((2 * n, 2 * n))
    for r in range(2 * n):
        for c in range(r + 1):
            i, j, p, q = r >> 1, c >> 1, r & 1, c & 1
            a = A[i, j] if uplo == "L" else np.conj(A[j, i])
            if i == j:
                a = a.real + 0j
            R[r, c] = a.real if p == q else (a.imag if p else -a.imag)
    return R


def fromreal(R, Rinv, A, uplo):
    """cplx_fromreal_kernel: the stored triangle of A and the uplo triangle
    of Tinv from the lower triangles of R and Rinv."""
    n = A.shape[0]
    A = A.copy()
    T = np.zeros((n, n), complex)
    for j in range(n):
        for i in range(j, n):
            e = (2 * i, 2 * j)
            l = complex(R[e], 0.0 if i == j else R[2 * i + 1, 2 * j])
            v = complex(Rinv[e], 0.0 if i == j else Rinv[2 * i + 1, 2 * j])
            if uplo == "L":
                A[i, j], T[i, j] = l, v
            else:
                A[j, i], T[j, i] = np.conj(l), np.conj(v)
    return A, T


def real_cholesky_lower(Rl):
    """Cholesky of the symmetric matrix whose lower triangle Rl holds;
    returns (L, 0) or (None, first failing order + 1)."""
    n = Rl.shape[0]
    M = np.tril(Rl).copy()
    for j in range(n):
        d = M[j, j] - M[j, :j] @ M[j, :j]
        if not d > 0:
            return None, j + 1
        M[j, j] = np.sqrt(d)
        for i in range(j + 1, n):
            M[i, j] = (M[i, j] - M[i, :j] @ M[j, :j]) / M[j, j]
    return np.tril(M), 0


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("unit", [False, True])
def test_trtri_leaf(uplo, unit):
    r = np.random.default_rng(3 + unit)
    n = 33
    # off-diagonal entries / n (the reference TRSM test's conditioning)
    A = (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))) / n + (1.5 + 0.5j) * np.eye(n)
    T = np.tril(A) if uplo == "L" else np.triu(A)
    if unit:
        np.fill_diagonal(T, 1.0)
    inv = trtri_leaf(A, uplo, unit)
    assert np.abs(inv @ T - np.eye(n)).max() < 1e-13


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_realform_potrf(uplo):
    r = np.random.default_rng(5)
    n = 24
    G = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    H = G @ G.conj().T + n * np.eye(n)
    # the stored triangle only; the other one and the diagonal's imaginary
    # parts hold junk that must not be read
    S = np.tril(H) if uplo == "L" else np.triu(H)
    S = S + (np.triu(np.ones((n, n)), 1) if uplo == "L" else np.tril(np.ones((n, n)), -1)) * (7 - 3j)
    S[np.diag_indices(n)] += 0.25j
    Rl = realform(S, uplo)
    Lr, info = real_cholesky_lower(Rl)
    assert info == 0
    A, T = fromreal(Lr, np.linalg.inv(Lr), S, uplo)
    if uplo == "L":
        F = np.tril(A)
        assert np.abs(F @ F.conj().T - H).max() / np.abs(H).max() < 1e-14
        assert np.abs(T @ F - np.eye(n)).max() < 1e-13
        assert np.array_equal(np.triu(A, 1), np.triu(S, 1))  # other triangle untouched
    else:
        F = np.triu(A)
        assert np.abs(F.conj().T @ F - H).max() / np.abs(H).max() < 1e-14
        assert np.abs(T @ F - np.eye(n)).max() < 1e-13
        assert np.array_equal(np.tril(A, -1), np.tril(S, -1))
    assert np.all(np.diag(A).imag == 0)
    # a failing complex pivot q fails the real form at order 2q: the leaf's
    # pair mode reports (2 base + 2q) / 2 + 1 = base + q + 1
    H2 = H.copy()
    H2[17, 17] = -1.0
    S2 = np.tril(H2) if uplo == "L" else np.triu(H2)
    _, info2 = real_cholesky_lower(realform(S2, uplo))
    base = 5
    assert info2 == 2 * 17 + 1
    assert (2 * base + info2 - 1) // 2 + 1 == base + 17 + 1
