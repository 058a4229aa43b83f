"""Seeded inputs shared by the tests (NumPy PCG64 streams, Fortran order)."""
import numpy as np


def rng(seed):
    return np.random.default_rng(seed)


def uniform(r, m, n, lo=-0.5, hi=0.5, dtype=np.float64):
    return np.asfortranarray(r.uniform(lo, hi, size=(m, n)).astype(dtype))


def small_int(r, m, n, lo=-8, hi=8, dtype=np.float64):
    return np.asfortranarray(r.integers(lo, hi + 1, size=(m, n)).astype(dtype))


def spd(r, n, dtype=np.float64):
    """A = G G^T + n I with G ~ U[-1,1] (the benchmark's SPD family)."""
    G = r.uniform(-1.0, 1.0, size=(n, n))
    A = G @ G.T + n * np.eye(n)
    return np.asfortranarray(A.astype(dtype))


def tri_matrix(r, n, uplo_lower, diag_unit, dtype=np.float64):
    """The reference TRSM test matrix (testing/ztesting_trsm.c:129-134):
    uniform entries scaled by 1/n with a unit diagonal."""
    A = r.uniform(0.0, 1.0, size=(n, n)) / n
    np.fill_diagonal(A, 1.0)
    return np.asfortranarray(A.astype(dtype))


def refined_solution(A, L, B, X):
    """X + A^-1 (B - A X) with the residual evaluated in 80-bit long double
    (its error ~ 2^-64 sqrt(n) |A| |X|) and the correction solved with the
    lower Cholesky factor L in FP64: a reference solution of A X = B whose
    error is far below FP64 rounding for a well-conditioned A.  Long-double
    products are slow: pass only the columns to be checked."""
    from scipy.linalg import solve_triangular
    R = B.astype(np.longdouble) - A.astype(np.longdouble) @ X.astype(np.longdouble)
    D = solve_triangular(L, np.asarray(R, dtype=np.float64), lower=True, check_finite=False)
    D = solve_triangular(L, D, lower=True, trans=1, check_finite=False)
    return np.asarray(X.astype(np.longdouble) + D.astype(np.longdouble), dtype=np.float64)
