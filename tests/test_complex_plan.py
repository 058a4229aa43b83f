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
