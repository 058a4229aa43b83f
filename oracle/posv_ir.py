"""CPU restatement of the mixed-precision SPD solve's refinement loop
(test infrastructure only).

The library's loop (xkblas_dposv_ir) follows LAPACK dsposv's scheme: X0 =
F^-1 B with a low-precision Cholesky factor F, then corrections d = F^-1 r,
r = B - A X in FP64, at most itermax corrections, after which (or after a
failed low-precision factor) the accurate factor is used.  Its stopping rule,
per column j of x_i = x_{i-1} + d_i:

* dsposv's backward test on the residual the correction was computed from,
  ||r||_inf <= sqrt(n) ||x||_inf ||A||_inf eps (eps = 2^-53; dsposv continues
  while rnrm > xnrm * cte);
* a forward test: ||d_i|| <= tau ||x||, or from the second correction on,
  with rho = ||d_i|| / ||d_{i-1}||, rho > 1/2 (no further progress) or
  rho / (1 - rho) ||d_i|| <= tau ||x|| (the error left in x_i); tau = 2^-48.

At the correction limit only the backward test is applied.  iter = the number
of corrections, or -(itermax + 1) when the limit is reached without passing.
"""
import numpy as np
from scipy.linalg import solve_triangular

EPS = 2.0 ** -53
TAU = 2.0 ** -48


def column_done(res_ok, corr, dn, dprev, xn, tau=TAU):
    """The per-column stopping test after correction number `corr`."""
    if not res_ok:
        return False
    if dn <= tau * xn:
        return True
    if corr < 2 or not dprev > 0.0:
        return False
    rho = dn / dprev
    if rho > 0.5:
        return True
    return rho / (1.0 - rho) * dn <= tau * xn


def refine(A, B, F, itermax=30, forward=True, tau=TAU):
    """X and the iteration code for A X = B with the lower Cholesky factor F
    standing in for the low-precision one.  forward=False applies dsposv's
    backward test alone (a correction is then taken only while it fails)."""
    n = A.shape[0]
    solve = lambda Y: solve_triangular(F.T, solve_triangular(F, Y, lower=True), lower=False)
    anrm = np.abs(A).sum(axis=1).max()
    cte = anrm * EPS * np.sqrt(n)
    X = solve(B)
    nrhs = B.shape[1]
    dprev = np.zeros(nrhs)
    for it in range(itermax + 1):
        R = B - A @ X
        rn = np.abs(R).max(axis=0)
        xn = np.abs(X).max(axis=0)
        res_ok = rn <= xn * cte
        if it == itermax or (not forward and np.all(res_ok)):
            if np.all(res_ok):
                return X, it
            break
        D = solve(R)
        dn = np.abs(D).max(axis=0)
        X = X + D
        if forward and all(column_done(res_ok[j], it + 1, dn[j], dprev[j], xn[j], tau) for j in range(nrhs)):
            return X, it + 1
        dprev = dn
    return X, -(itermax + 1)
