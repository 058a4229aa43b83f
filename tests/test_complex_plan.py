"""CPU check of the complex-GEMM plan used by dev_cgemm (xkblas_amd/csrc/
complex.hip): the real block operands exactly as cplx_pack_a / cplx_pack_b
lay them out, one real product with the same transpose flags, and
cplx_combine's alpha/beta fold reproduce the complex product for every
op(A) x op(B) in {N, T, C}^2.  A wrong sign of an imaginary block, a swapped
block or a wrong transpose flag fails here."""
import itertools

import numpy as np
import pytest

N_, T_, C_ = 111, 112, 113


def pack_a(A, sigma):
    """Y = [Ar, -sigma Ai; sigma Ai, Ar] (cplx_pack_a_kernel)."""
    Ar, Ai = A.real, A.imag
    return np.block([[Ar, -sigma * Ai], [sigma * Ai, Ar]])


def pack_b(B, vertical, sign):
    """[Br; Bi] or [Br, sign Bi] (cplx_pack_b_kernel)."""
    return np.vstack([B.real, B.imag]) if vertical else np.hstack([B.real, sign * B.imag])


def plan(ta, tb, A, B, alpha, beta, C):
    YA = pack_a(A, -1 if ta == T_ else 1)
    YB = pack_b(B, tb == N_, -1 if tb == C_ else 1)
    opA = YA.T if ta != N_ else YA
    opB = YB.T if tb != N_ else YB
    P = opA @ opB
    m = C.shape[0]
    return alpha * (P[:m] + 1j * P[m:]) + beta * C


@pytest.mark.parametrize("ta,tb", list(itertools.product([N_, T_, C_], repeat=2)))
def test_block_plan_matches_complex_product(ta, tb):
    r = np.random.default_rng(ta * 3 + tb)
    m, n, k = 7, 5, 6
    c = lambda p, q: r.standard_normal((p, q)) + 1j * r.standard_normal((p, q))
    A = c(m, k) if ta == N_ else c(k, m)
    B = c(k, n) if tb == N_ else c(n, k)
    C = c(m, n)
    op = lambda X, t: X if t == N_ else (X.T if t == T_ else X.conj().T)
    ref = (0.5 - 2j) * (op(A, ta) @ op(B, tb)) + (1 + 1j) * C
    np.testing.assert_allclose(plan(ta, tb, A, B, 0.5 - 2j, 1 + 1j, C), ref, atol=1e-12)


def pack2(B, vertical, swap, sr, si):
    """[sr Br, si Bi] side by side, or stacked when vertical; swap puts
    si Bi first (cplx_pack2_kernel)."""
    first, second = (si * B.imag, sr * B.real) if swap else (sr * B.real, si * B.imag)
    return np.vstack([first, second]) if vertical else np.hstack([first, second])


def triangle_plan(ta, tb, A, B, n, k):
    """The triangle path of cgemm_impl: U = [Xr Xi], V1 = [Zr -Zi],
    V2 = [Zi Zr] with X = op(A), Z = op(B)^T, packed from the stored
    matrices with the same flags; returns (Pr, Pi, same)."""
    ta_t = ta != N_
    tb_t = tb == N_
    zi = -1 if tb == C_ else 1
    U = pack2(A, ta_t, 0, 1, -1 if ta == C_ else 1)
    V1 = pack2(B, tb_t, 0, 1, -zi)
    V2 = pack2(B, tb_t, 1, 1, zi)
    same = A is B and ta_t == tb_t and (-1 if ta == C_ else 1) == -zi
    opu = U.T if ta_t else U
    opv = lambda V: V if tb_t else V.T
    return opu @ opv(V1), opu @ opv(V2), same, np.array_equal(U, V1)


@pytest.mark.parametrize("ta,tb", list(itertools.product([N_, T_, C_], repeat=2)))
def test_triangle_plan_matches_complex_product(ta, tb):
    r = np.random.default_rng(ta * 5 + tb)
    n, k = 6, 4
    c = lambda p, q: r.standard_normal((p, q)) + 1j * r.standard_normal((p, q))
    A = c(n, k) if ta == N_ else c(k, n)
    B = c(k, n) if tb == N_ else c(n, k)
    op = lambda X, t: X if t == N_ else (X.T if t == T_ else X.conj().T)
    P = op(A, ta) @ op(B, tb)
    Pr, Pi, _, _ = triangle_plan(ta, tb, A, B, n, k)
    np.testing.assert_allclose(Pr, P.real, atol=1e-12)
    np.testing.assert_allclose(Pi, P.imag, atol=1e-12)


@pytest.mark.parametrize("trans", [N_, C_])
def test_triangle_plan_herk_real_part_is_a_syrk(trans):
    """HERK (B = A, op(B) = op(A)^H): the real part's operands are one packed
    matrix, so the flag that reuses U for V1 is exactly when they are equal."""
    r = np.random.default_rng(trans)
    n, k = 5, 3
    A = r.standard_normal((n, k) if trans == N_ else (k, n)) + 1j * r.standard_normal((n, k) if trans == N_ else (k, n))
    ta, tb = (N_, C_) if trans == N_ else (C_, N_)
    Pr, Pi, same, equal = triangle_plan(ta, tb, A, A, n, k)
    assert same and equal
    X = A if trans == N_ else A.conj().T
    np.testing.assert_allclose(Pr + 1j * Pi, X @ X.conj().T, atol=1e-12)


def rho_form(A):
    """The 2n x 2n real form [rho(a_ij)], rho(z) = [re -im; im re]
    (cplx_realform_kernel, full matrix instead of its lower triangle)."""
    n = A.shape[0]
    M = np.empty((2 * n, 2 * n))
    M[0::2, 0::2] = A.real
    M[1::2, 1::2] = A.real
    M[1::2, 0::2] = A.imag
    M[0::2, 1::2] = -A.imag
    return M


def test_real_form_cholesky_is_rho_of_complex_factor():
    """The complex POTRF's diagonal blocks: the real Cholesky factor of the
    real form is rho(L) and its inverse rho(inv(L)), so that cplx_fromreal's
    l_ij = R[2i][2j] + i R[2i+1][2j] reads L and inv(L).  A sign slip in rho
    (a transposed imaginary block) makes the real form non-symmetric or the
    read-back the conjugate factor, which fails here."""
    r = np.random.default_rng(41)
    n = 9
    G = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    A = G @ G.conj().T + n * np.eye(n)
    A = (A + A.conj().T) / 2
    np.fill_diagonal(A, A.diagonal().real)
    L = np.linalg.cholesky(A)
    M = rho_form(A)
    np.testing.assert_allclose(M, M.T, atol=0)
    R = np.linalg.cholesky(M)
    Rinv = np.linalg.inv(R)
    back = lambda X: X[0::2, 0::2] + 1j * X[1::2, 0::2]
    np.testing.assert_allclose(np.tril(back(R)), L, atol=1e-12)
    np.testing.assert_allclose(np.tril(back(Rinv)), np.linalg.inv(L), atol=1e-12)
    np.testing.assert_allclose(R, rho_form(L), atol=1e-12)
    # the diagonal blocks of the real factor are l_ii I: the odd rows' entry
    # left of the diagonal is zero
    np.testing.assert_allclose(np.diag(R, -1)[0::2], 0, atol=1e-12)
