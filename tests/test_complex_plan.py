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
