"""CPU restatement of dev_potrf's trailing-update schedule (xkblas_amd/csrc/
blas3.cpp): one rank-NB update per panel, or paired updates (an even step
updates only the next block column; the odd step after it applies both
panels as one rank-2NB product, the first ~29 % of the trailing columns at
once and the rest at the next step).  The panel is formed in two parts: its
rows in the next block row first (all the lookahead reads), then the rest.  Each step's products are applied in the
same order and on the same blocks as the library's launches, with NumPy's
Cholesky for the diagonal blocks; the result must be the Cholesky factor of A
for both triangles, ragged last blocks and odd / even block counts.  A part
applied twice or dropped, or a wrong column split, fails."""
import numpy as np
import pytest


def sim(n, nb, agg, uplo):
    r = np.random.default_rng(n * 7 + nb)
    G = r.standard_normal((n, n)); S = G @ G.T + n * np.eye(n)
    A = S.copy()
    nt = (n + nb - 1) // nb
    tsz = lambda i: n - i * nb if i == nt - 1 else nb
    def sym(M, lower):
        M = np.tril(M) if lower else np.triu(M)
        return M + M.T - np.diag(np.diag(M))
    def diag(k):
        b = slice(k*nb, k*nb+tsz(k))
        if uplo == "L":
            L = np.linalg.cholesky(sym(A[b, b], True)); A[b, b] = L; return np.linalg.inv(L)
        U = np.linalg.cholesky(sym(A[b, b], False)).T; A[b, b] = U; return np.linalg.inv(U)
    def pair_split(rr):
        w = int((0.2929 * rr) / nb + 0.5) * nb
        return min(rr, max(nb, w))
    inv = {0: diag(0)}
    pend = None
    for k in range(nt):
        kk = tsz(k)
        if k == nt - 1: break
        rest = n - (k+1)*nb
        pair = agg and (k & 1) == 1
        prev, pend = pend, None
        k1 = tsz(k+1); r2 = rest - k1
        c0 = k - 1 if pair else k
        r0 = (k+1)*nb
        if uplo == "L":
            # the panel's rows in block row k+1, the lookahead, then the rest
            A[r0:r0+k1, k*nb:k*nb+kk] = A[r0:r0+k1, k*nb:k*nb+kk] @ inv[k].T
            U1 = A[r0:r0+k1, c0*nb:(k+1)*nb].copy()
            A[r0:r0+k1, r0:r0+k1] -= np.tril(U1 @ U1.T)
            A[r0+k1:, k*nb:k*nb+kk] = A[r0+k1:, k*nb:k*nb+kk] @ inv[k].T
            U = A[r0:, c0*nb:(k+1)*nb].copy()
            if r2 > 0:
                A[r0+k1:, r0:r0+k1] -= U[k1:] @ U[:k1].T
                c = r0 + k1
                if not agg:
                    A[c:, c:] -= np.tril(U[k1:] @ U[k1:].T)
                elif pair:
                    w1 = pair_split(r2)
                    A[c:c+w1, c:c+w1] -= np.tril(U[k1:k1+w1] @ U[k1:k1+w1].T)
                    if w1 < r2:
                        A[c+w1:, c:c+w1] -= U[k1+w1:] @ U[k1:k1+w1].T
                        pend = (U[k1+w1:], c + w1)
            if prev is not None:
                Up, cp = prev
                A[cp:, cp:] -= np.tril(Up @ Up.T)
        else:
            A[k*nb:k*nb+kk, r0:r0+k1] = inv[k].T @ A[k*nb:k*nb+kk, r0:r0+k1]
            U1 = A[c0*nb:(k+1)*nb, r0:r0+k1].copy()
            A[r0:r0+k1, r0:r0+k1] -= np.triu(U1.T @ U1)
            A[k*nb:k*nb+kk, r0+k1:] = inv[k].T @ A[k*nb:k*nb+kk, r0+k1:]
            U = A[c0*nb:(k+1)*nb, r0:].copy()
            if r2 > 0:
                A[r0:r0+k1, r0+k1:] -= U[:, :k1].T @ U[:, k1:]
                c = r0 + k1
                if not agg:
                    A[c:, c:] -= np.triu(U[:, k1:].T @ U[:, k1:])
                elif pair:
                    w1 = pair_split(r2)
                    A[c:c+w1, c:c+w1] -= np.triu(U[:, k1:k1+w1].T @ U[:, k1:k1+w1])
                    if w1 < r2:
                        A[c:c+w1, c+w1:] -= U[:, k1:k1+w1].T @ U[:, k1+w1:]
                        pend = (U[:, k1+w1:], c + w1)
            if prev is not None:
                Up, cp = prev
                A[cp:, cp:] -= np.triu(Up.T @ Up)
        inv[k+1] = diag(k+1)
    assert pend is None  # every deferred part was applied
    if uplo == "L":
        Lf = np.tril(A); return np.abs(Lf @ Lf.T - S).max() / np.abs(S).max()
    Uf = np.triu(A); return np.abs(Uf.T @ Uf - S).max() / np.abs(S).max()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("agg", [False, True])
@pytest.mark.parametrize("n,nb", [(5, 4), (16, 4), (17, 4), (23, 5), (40, 8), (41, 8), (65, 8), (90, 5)])
def test_potrf_schedule_gives_the_factor(n, nb, agg, uplo):
    assert sim(n, nb, agg, uplo) < 1e-12
