"""Mixed-precision SPD solve with FP64 iterative refinement (xkblas_dposv_ir).

The stopping rule combines LAPACK dsposv's backward test (every column of
r = B - A X has ||r||_inf <= sqrt(n) ||x||_inf ||A||_inf eps, eps = 2^-53)
with a forward test on the corrections (restated and pinned on the CPU in
oracle/posv_ir.py); the residuals sum short product runs without error, so X
converges to within 1e-14 (max-abs relative) of the exact solution.  The
tests check dsposv's test on the host from the returned X (FP64 NumPy
residual, a factor 2 of slack for the different summation order), compare X
at 1e-14 with LAPACK's FP64 solve refined once with a long-double residual
(an FP64 solve alone is only accurate to ~1e-14 at n = 4096), and up to
n = 2048 with the oracle's tiled Cholesky + two tiled triangular solves,
and with LAPACK's FP64 solve for the fallbacks, check that only the uplo triangle of
A is read (the other one holds NaN), and force both fallbacks of dsposv: the
correction limit (itermax 0) and a matrix that is not positive definite.
The solves with the reduced factor run in FP32 within a range of ||A||: a
scaled family checks both sides of it.
"""
import numpy as np
import pytest

from helpers import rng, spd, uniform, refined_solution

pytestmark = pytest.mark.gpu

LOWER, UPPER = 122, 121
EPS = 2.0 ** -53


def _stored(A, uplo):
    """A with the triangle that is not read set to NaN."""
    S = A.copy(order="F")
    iu = np.triu_indices(A.shape[0], 1) if uplo == LOWER else np.tril_indices(A.shape[0], -1)
    S[iu] = np.nan
    return S


def _criterion(A, B, X):
    n = A.shape[0]
    R = B - A @ X
    rn = np.abs(R).max(axis=0)
    xn = np.abs(X).max(axis=0)
    anrm = np.abs(A).sum(axis=1).max()
    return rn / (xn * anrm * EPS * np.sqrt(n))


def _oracle_solve(orc, uplo, A, B, nb):
    """X from the oracle's tiled Cholesky and two tiled triangular solves."""
    F = A.copy(order="F")
    assert orc.dpotrf_tiled(uplo, F, nb) == 0
    X = B.copy(order="F")
    first, second = (111, 112) if uplo == LOWER else (112, 111)
    orc.dtrsm_tiled(141, uplo, first, 131, 1.0, F, X, nb)
    orc.dtrsm_tiled(141, uplo, second, 131, 1.0, F, X, nb)
    return X


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n,nrhs,nb", [(300, 1, 128), (1000, 37, 256), (2048, 64, 512), (1100, 5, 1100),
                                       (4096, 64, 1024)])
def test_dposv_ir_converges(xk, orc, uplo, n, nrhs, nb):
    L = xk.lib()
    r = rng(n * 7 + nrhs)
    A = spd(r, n)
    B = uniform(r, n, nrhs)
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(nb, 0)
    try:
        X, info, it = xk.posv_ir(uplo, _stored(A, uplo), B)
    finally:
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
    assert info == 0
    if nb >= n:
        # one diagonal block: the whole factor is FP64; one correction of
        # rounding size shows it
        assert 1 <= it <= 2, it
    else:
        # split products of the reduced factor are not FP64-accurate: the
        # first correction is large, the forward test needs a second
        assert 2 <= it <= 10, it
    assert np.all(_criterion(A, B, X) < 2.0)
    # forward error on up to 4 columns against LAPACK's FP64 solution refined
    # once with an extended-precision residual (an FP64 solve alone, the
    # oracle's tiled one included, is only accurate to ~1e-14 at n = 4096)
    Lc = np.linalg.cholesky(A)
    cols = np.arange(min(nrhs, 4))
    Xl = np.linalg.solve(A, B[:, cols])
    Xs = refined_solution(A, Lc, B[:, cols], Xl)
    assert np.abs(X[:, cols] - Xs).max() / np.abs(Xs).max() <= 1e-14
    if n <= 2048:
        Xr = _oracle_solve(orc, uplo, A, B, nb)
        assert np.abs(Xr[:, cols] - Xs).max() / np.abs(Xs).max() <= 1e-13
        assert np.abs(X - Xr).max() / np.abs(Xr).max() <= 2e-14


@pytest.mark.parametrize("mode", [0, 2])
def test_dposv_ir_fallback_after_itermax(xk, mode):
    """itermax 0: no correction allowed, so X comes from the factor of the
    current mode (FP64 MFMA or split-BF16 at its default count) and
    *iter = -1."""
    L = xk.lib()
    r = rng(91 + mode)
    n, nrhs = 700, 9
    A = spd(r, n)
    B = uniform(r, n, nrhs)
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(128, 0)
    L.xkblas_set_posv_ir(7, 0)
    L.xkblas_set_modemath(mode)
    try:
        X, info, it = xk.posv_ir(LOWER, _stored(A, LOWER), B)
    finally:
        L.xkblas_set_posv_ir(0, -1)
        L.xkblas_set_modemath(0)
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
    assert (info, it) == (0, -1)
    assert np.all(_criterion(A, B, X) < 2.0)
    Xr = np.linalg.solve(A, B)
    assert np.abs(X - Xr).max() / np.abs(Xr).max() <= 1e-14


def test_dposv_ir_more_moduli_fewer_corrections(xk):
    """More CRT moduli make the reduced factor more accurate: never more
    corrections."""
    L = xk.lib()
    r = rng(93)
    n = 1024
    A = spd(r, n)
    B = uniform(r, n, 3)
    its = []
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(256, 0)
    try:
        for m in (6, 7, 9):
            L.xkblas_set_posv_ir(m, 30)
            X, info, it = xk.posv_ir(UPPER, _stored(A, UPPER), B)
            xk.invalidate()
            assert info == 0 and it >= 0
            assert np.all(_criterion(A, B, X) < 2.0)
            its.append(it)
    finally:
        L.xkblas_set_posv_ir(0, -1)
        L.xkblas_set_param(nb0, 0)
    assert its[0] >= its[1] >= its[2] and its[0] >= 1, its


@pytest.mark.parametrize("moduli", [5, 7])
def test_dposv_ir_coarse_factor(xk, moduli):
    """A coarser reduced factor (5 CRT moduli, about 2^-16 per product) still
    converges on a well-conditioned A, with more corrections than the default
    7, to the same 1e-14 bar against the refined solution."""
    L = xk.lib()
    r = rng(103)
    n, nrhs = 2048, 8
    A = spd(r, n)
    B = uniform(r, n, nrhs)
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(512, 0)
    L.xkblas_set_posv_ir(moduli, 30)
    try:
        X, info, it = xk.posv_ir(LOWER, _stored(A, LOWER), B)
    finally:
        L.xkblas_set_posv_ir(0, -1)
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
    assert info == 0 and 2 <= it <= 6, (info, it)
    cols = np.arange(4)
    Lc = np.linalg.cholesky(A)
    Xs = refined_solution(A, Lc, B[:, cols], np.linalg.solve(A, B[:, cols]))
    assert np.abs(X[:, cols] - Xs).max() / np.abs(Xs).max() <= 1e-14
    assert np.all(_criterion(A, B, X) < 2.0)


@pytest.mark.parametrize("s", [-120, -60, 80, 95])
def test_dposv_ir_scaled_matrix(xk, s):
    """The solves with the reduced factor run in FP32 while ||A||_inf stays in
    [2^-100, 2^100] (||A0||_inf ~ 2^13 here: s = -60 and 80 inside, -120 and
    95 outside, where they run in FP64), the right-hand sides scaled per
    column by a power of two.  A = 2^s A0, B = 2^(s // 2) B0: X is A0^-1 B0
    scaled exactly, at the 1e-14 bar either way."""
    L = xk.lib()
    r = rng(101)
    n, nrhs = 1000, 6
    A0 = spd(r, n)
    B0 = uniform(r, n, nrhs)
    A = np.asfortranarray(np.ldexp(A0, s))
    B = np.asfortranarray(np.ldexp(B0, s // 2))
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(256, 0)
    try:
        X, info, it = xk.posv_ir(LOWER, _stored(A, LOWER), B)
    finally:
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
    assert info == 0 and 2 <= it <= 10, (info, it)
    cols = np.arange(4)
    Lc = np.linalg.cholesky(A0)
    Xs = refined_solution(A0, Lc, B0[:, cols], np.linalg.solve(A0, B0[:, cols]))
    Xu = np.ldexp(X[:, cols], s - s // 2)
    assert np.abs(Xu - Xs).max() / np.abs(Xs).max() <= 1e-14


@pytest.mark.parametrize("bad", [0, 333, 999])
def test_dposv_ir_not_positive_definite(xk, bad):
    """A negative pivot at order bad+1: the reduced factor fails (*iter = -3)
    and so does the accurate one, whose LAPACK info is returned."""
    L = xk.lib()
    r = rng(97)
    n = 1000
    A = np.asfortranarray(np.diag(r.uniform(1.0, 2.0, n)))
    A[bad, bad] = -1.0
    B = uniform(r, n, 2)
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(256, 0)
    try:
        X, info, it = xk.posv_ir(LOWER, A, B)
    finally:
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
    assert (info, it) == (bad + 1, -3)
