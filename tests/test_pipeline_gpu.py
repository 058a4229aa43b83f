"""Row-panel pipelining of products whose operands start in host memory.

A DGEMM whose op(A) or C is not yet on the device runs one tile row of C at
a time, each panel waiting only for its own copies, and the copy back of C
follows the panels.  Every element sees the same operands and inner order
as the single launch of the resident path, so the results must be equal bit
for bit, in FP64 MFMA and in the split modes (row scalings span the whole
inner dimension; B's residue planes are computed once), for ragged shapes,
every transpose pair and beta = 0 (C not read) or not.
"""
import numpy as np
import pytest

from helpers import rng, uniform

pytestmark = pytest.mark.gpu

N_, T_ = 111, 112


def _run(xk, pipeline, ta, tb, alpha, A, B, beta, C0, nb):
    L = xk.lib()
    xk.invalidate()
    L.xkblas_set_param(nb, 0)
    L.xkblas_set_pipeline(pipeline)
    try:
        C = C0.copy(order="F")
        assert xk.gemm(ta, tb, alpha, A, B, beta, C) == 0
    finally:
        L.xkblas_set_pipeline(1)
        xk.invalidate()
    return C


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("ta,tb", [(N_, N_), (T_, N_), (N_, T_), (T_, T_)])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_pipelined_equals_one_launch(xk, orc, mode, ta, tb, beta):
    M, N, K, nb = 700, 300, 500, 128
    r = rng(31 + ta + 2 * tb + (beta != 0) + 7 * mode)
    A = uniform(r, M, K) if ta == N_ else uniform(r, K, M)
    B = uniform(r, K, N) if tb == N_ else uniform(r, N, K)
    C0 = uniform(r, M, N)
    L = xk.lib()
    L.xkblas_set_modemath(mode)
    try:
        Cp = _run(xk, 1, ta, tb, 1.5, A, B, beta, C0, nb)
        Cr = _run(xk, 0, ta, tb, 1.5, A, B, beta, C0, nb)
    finally:
        L.xkblas_set_modemath(0)
    assert np.array_equal(Cp, Cr)
    Cref = orc.dgemm_tiled(ta, tb, 1.5, A, B, beta, C0.copy(order="F"), nb)
    assert orc.gemm_residual(1.5, A, B, beta, C0, Cref, Cp) <= 3.0


def test_pipelined_then_resident_reuse(xk):
    """A pipelined product leaves its operands resident: the same call again
    (one launch, nothing to fetch) gives the same bits, and a partial copy
    back of C (rows of two panels) waits only for the panels it reads."""
    M, N, K, nb = 1000, 260, 300, 256
    r = rng(41)
    A, B, C0 = uniform(r, M, K), uniform(r, K, N), uniform(r, M, N)
    L = xk.lib()
    xk.invalidate()
    L.xkblas_set_param(nb, 0)
    C = C0.copy(order="F")
    Cs = C0.copy(order="F")
    xk.gemm(N_, N_, 1.0, A, B, 0.0, C, sync=False)
    L.xkblas_memory_coherent_async(0, 0, 300, N, C.ctypes.data, M, 8)
    L.xkblas_sync()
    top = C[:300].copy()
    xk.coherent(C)
    xk.gemm(N_, N_, 1.0, A, B, 0.0, Cs)
    xk.invalidate()
    assert np.array_equal(top, C[:300])
    assert np.array_equal(C, Cs)
    np.testing.assert_allclose(C, A @ B, rtol=0, atol=1e-12)
