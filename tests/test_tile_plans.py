"""CPU checks of the workgroup -> tile plans of the GEMM-shaped kernels:

* tri_tile<GM> (xkblas_amd/csrc/tiles.h): the XCD-blocked, row-grouped order
  of the lower-triangle tiles is a bijection onto {(i, j): 0 <= j <= i < T}
  for every order up to 70 and around 128/256/512 tiles, GM = 4 and 8 (a
  wrong group size, a dropped diagonal tile or an off-by-one in the quadratic
  guess fails);
* the per-tile inner ranges of a triangular operand (KRANGE_* in
  xkblas_amd/csrc/gemm.hip, used by TRMM) keep every non-zero term: the
  restricted tile sums equal the full product for op(A) lower / upper and
  op(B) upper / lower, ragged sizes, tile orders 32/64/128.
"""
import numpy as np
import pytest


def tri_tile(pid, total, T, GM):
    """Python restatement of tri_tile<GM> (float32 first guess, as on the GPU)."""
    if total % 8 == 0:
        pid = (pid & 7) * (total >> 3) + (pid >> 3)
    before = lambda h: (GM * GM * h * (h - 1)) // 2 + h * (GM * (GM + 1)) // 2
    a = np.float32(0.5 * GM * GM)
    b = np.float32(0.5 * GM * (GM + 1)) - a
    g = int((-b + np.sqrt(np.float32(b * b + np.float32(4.0) * a * np.float32(pid)))) / (np.float32(2.0) * a))
    while g > 0 and before(g) > pid:
        g -= 1
    while before(g + 1) <= pid:
        g += 1
    r = pid - before(g)
    r0 = g * GM
    gr = min(GM, T - r0)
    if r < r0 * gr:
        return r0 + r % gr, r // gr
    rr, jj = r - r0 * gr, 0
    while rr >= gr - jj:
        rr -= gr - jj
        jj += 1
    return r0 + jj + rr, r0 + jj


@pytest.mark.parametrize("GM", [4, 8])
def test_tri_tile_is_a_bijection(GM):
    for T in list(range(1, 71)) + [127, 128, 129, 255, 256, 257, 511, 512]:
        total = T * (T + 1) // 2
        seen = set()
        for pid in range(total):
            i, j = tri_tile(pid, total, T, GM)
            assert 0 <= j <= i < T, (GM, T, pid, i, j)
            seen.add((i, j))
        assert len(seen) == total, (GM, T)


KRANGE_A_LOWER, KRANGE_A_UPPER, KRANGE_B_UPPER, KRANGE_B_LOWER = 1, 2, 3, 4


def krange_product(opA, opB, kmode, BM):
    """C = opA opB computed tile by tile over each tile's inner range."""
    M, K = opA.shape
    N = opB.shape[1]
    C = np.zeros((M, N))
    for m0 in range(0, M, BM):
        for n0 in range(0, N, BM):
            kbeg, kend = 0, K
            if kmode == KRANGE_A_LOWER:
                kend = min(K, m0 + BM)
            elif kmode == KRANGE_A_UPPER:
                kbeg = m0
            elif kmode == KRANGE_B_UPPER:
                kend = min(K, n0 + BM)
            elif kmode == KRANGE_B_LOWER:
                kbeg = n0
            # the kernel walks whole 16-deep slices from kbeg // 16
            kbeg = kbeg // 16 * 16
            C[m0:m0 + BM, n0:n0 + BM] = opA[m0:m0 + BM, kbeg:kend] @ opB[kbeg:kend, n0:n0 + BM]
    return C


@pytest.mark.parametrize("BM", [32, 64, 128])
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("left", [True, False])
def test_krange_keeps_every_nonzero_term(BM, lower, left):
    r = np.random.default_rng(BM + 2 * lower + left)
    n, other = 333, 170
    T = np.tril(r.standard_normal((n, n))) if lower else np.triu(r.standard_normal((n, n)))
    X = r.standard_normal((n, other)) if left else r.standard_normal((other, n))
    if left:
        kmode = KRANGE_A_LOWER if lower else KRANGE_A_UPPER
        got, ref = krange_product(T, X, kmode, BM), T @ X
    else:
        kmode = KRANGE_B_LOWER if lower else KRANGE_B_UPPER
        got, ref = krange_product(X, T, kmode, BM), X @ T
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10)
