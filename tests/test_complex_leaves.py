"""CPU restatements of the complex single-wave leaves (xkblas_amd/csrc/
complex.hip), lane by lane as the kernels run them:

* cplx_trtri_leaf_kernel: lane j forward-substitutes column j of inv(L) in the
  lower view (Upper: L = U^T, inv(U) = inv(L)^T), unit or not;
* cplx_potrf_leaf_kernel: right-looking by columns, lane i owning row i of
  the trailing update, in the lower view (Upper: L = U^H, written back as
  U = L^H); a non-positive pivot stops at its column.

Each must give the inverse / the Hermitian factor of the textbook definition
(a conjugate missing in the update or the view, or a wrong loop bound, fails).
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


def potrf_leaf(A, uplo):
    n = A.shape[0]
    M = np.zeros((n, n), complex)
    for i in range(n):
        for k in range(i + 1):
            M[i, k] = A[i, k] if uplo == "L" else np.conj(A[k, i])
    for j in range(n):
        d = M[j, j].real
        if not d > 0:
            return None, j + 1
        sd = np.sqrt(d)
        M[j, j] = sd
        M[j + 1:, j] /= sd
        for i in range(j + 1, n):  # lane i
            for k in range(j + 1, i + 1):
                M[i, k] -= M[i, j] * np.conj(M[k, j])
    L = np.tril(M)
    return (L if uplo == "L" else L.conj().T), 0


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
def test_potrf_leaf(uplo):
    r = np.random.default_rng(5)
    n = 40
    G = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    H = G @ G.conj().T + n * np.eye(n)
    F, info = potrf_leaf(H, uplo)
    assert info == 0
    R = F @ F.conj().T if uplo == "L" else F.conj().T @ F
    assert np.abs(R - H).max() / np.abs(H).max() < 1e-14
    H2 = H.copy()
    H2[17, 17] = -1.0
    assert potrf_leaf(H2, uplo)[1] > 0
