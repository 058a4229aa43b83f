"""The refinement loop of the mixed-precision SPD solve, restated on the CPU
(oracle/posv_ir.py, the loop of dev_posv_ir in xkblas_amd/csrc/blas3.cpp):
a low-precision factor (the FP64 Cholesky factor rounded to `bits`
significant bits, standing in for the reduced-moduli split factor), X0 =
F^-1 B, corrections from FP64 residuals, dsposv's backward test plus a
forward test on the corrections.

Pinned by what the mathematics fixes: the solution reaches a relative error
<= 1e-14 against an accurate FP64 solve for every factor precision; the
exact factor needs at most one correction; a factor of more bits never needs
more corrections; no correction allowed gives dsposv's fallback code.  A
planted mistake (the forward test dropped: dsposv's backward test alone)
stops early with errors above 1e-14 on the same inputs."""
import numpy as np
import pytest

from oracle.posv_ir import EPS, column_done, refine


def rounded(X, bits):
    m, e = np.frexp(X)
    return np.ldexp(np.round(m * 2.0 ** bits) / 2.0 ** bits, e)


def problem(n, seed, nrhs=4):
    r = np.random.default_rng(seed)
    G = r.uniform(-1, 1, (n, n))
    A = G @ G.T + n * np.eye(n)
    B = r.uniform(-0.5, 0.5, (n, nrhs))
    return A, B


def relerr(X, Xs):
    return np.abs(X - Xs).max() / np.abs(Xs).max()


@pytest.mark.parametrize("n", [50, 200])
def test_refinement_converges_and_counts(n):
    A, B = problem(n, n, 3)
    L = np.linalg.cholesky(A)
    its = []
    for bits in (16, 24, 40):
        X, it = refine(A, B, rounded(L, bits))
        assert it >= 1
        R = B - A @ X
        assert np.all(np.abs(R).max(axis=0) <= np.abs(X).max(axis=0) * np.abs(A).sum(1).max() * EPS * np.sqrt(n))
        its.append(it)
    assert its[0] >= its[1] >= its[2]
    # the exact factor: one correction, of rounding size
    assert refine(A, B, L)[1] <= 1
    # no correction allowed: the dsposv fallback code
    assert refine(A, B, rounded(L, 16), itermax=0)[1] == -1


def test_forward_error_below_1e14():
    """n = 1500 (sqrt(n) eps = 4.3e-15 in dsposv's test), kappa ~ 2.3: every
    factor precision reaches <= 1e-14; dsposv's test alone does not."""
    n = 1500
    A, B = problem(n, n)
    L = np.linalg.cholesky(A)
    Xs = np.linalg.solve(A, B)
    worst_backward_only = 0.0
    for bits in (12, 16, 24, 32):
        F = rounded(L, bits)
        X, it = refine(A, B, F)
        assert it >= 1
        assert relerr(X, Xs) <= 1e-14, (bits, it, relerr(X, Xs))
        Xb, itb = refine(A, B, F, forward=False)
        assert itb <= it
        worst_backward_only = max(worst_backward_only, relerr(Xb, Xs))
    assert worst_backward_only > 1e-14


def test_column_rule_cases():
    # the backward test gates everything
    assert not column_done(False, 3, 0.0, 1.0, 1.0)
    # a negligible correction
    assert column_done(True, 1, 2.0 ** -49, 0.0, 1.0)
    assert not column_done(True, 1, 2.0 ** -47, 0.0, 1.0)
    # first correction: no contraction estimate yet
    assert not column_done(True, 1, 1e-7, 0.0, 1.0)
    # fast contraction: predicted error rho/(1-rho) |d| = 1e-21
    assert column_done(True, 2, 1e-14, 1e-7, 1.0)
    # slow contraction: 0.25/0.75 * 4e-14 > 2^-48
    assert not column_done(True, 2, 4e-14, 1.6e-13, 1.0)
    # no progress (rho > 1/2): the FP64 residual's rounding is reached
    assert column_done(True, 3, 1e-15, 1.5e-15, 1.0)
    # a zero solution column
    assert column_done(True, 1, 0.0, 0.0, 0.0)
