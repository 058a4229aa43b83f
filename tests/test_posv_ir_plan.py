"""CPU restatement of dev_posv_ir's refinement loop (xkblas_amd/csrc/blas3.cpp,
LAPACK dsposv's scheme): with a low-precision factor (here the FP64 Cholesky
factor rounded to `bits` significant bits, standing in for the reduced-moduli
split factor), X0 = F^-1 B, then r = B - A X, stop when every column has
||r||_inf < sqrt(n) ||x||_inf ||A||_inf eps (checked before each correction,
iteration 0 included), at most itermax corrections, else the code
-(itermax + 1).  Pins the stopping rule, the iteration count semantics and the
fallback code; a factor of more bits needs no more corrections."""
import numpy as np
import pytest
from scipy.linalg import solve_triangular

EPS = 2.0 ** -53


def rounded(X, bits):
    m, e = np.frexp(X)
    return np.ldexp(np.round(m * 2.0 ** bits) / 2.0 ** bits, e)


def refine(A, B, F, itermax=30):
    n = A.shape[0]
    solve = lambda Y: solve_triangular(F.T, solve_triangular(F, Y, lower=True), lower=False)
    anrm = np.abs(A).sum(axis=1).max()
    cte = anrm * EPS * np.sqrt(n)
    X = solve(B)
    for it in range(itermax + 1):
        R = B - A @ X
        if np.all(np.abs(R).max(axis=0) < np.abs(X).max(axis=0) * cte):
            return X, it
        if it == itermax:
            break
        X = X + solve(R)
    return X, -(itermax + 1)


@pytest.mark.parametrize("n", [50, 200])
def test_refinement_converges_and_counts(n):
    r = np.random.default_rng(n)
    G = r.uniform(-1, 1, (n, n))
    A = G @ G.T + n * np.eye(n)
    B = r.uniform(-0.5, 0.5, (n, 3))
    L = np.linalg.cholesky(A)
    its = []
    for bits in (16, 24, 40):
        X, it = refine(A, B, rounded(L, bits))
        assert it >= 1
        R = B - A @ X
        assert np.all(np.abs(R).max(axis=0) < np.abs(X).max(axis=0) * np.abs(A).sum(1).max() * EPS * np.sqrt(n))
        its.append(it)
    assert its[0] >= its[1] >= its[2]
    # the exact factor needs no correction (or one, from rounding)
    assert refine(A, B, L)[1] <= 1
    # no correction allowed: the dsposv fallback code
    assert refine(A, B, rounded(L, 16), itermax=0)[1] == -1
