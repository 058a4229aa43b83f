"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars (the reference test drivers' own normwise checks):
  GEMM  <= 3 eps   (testing/ztesting_gemm.c:288-306)
  SYRK  <= 10      (testing/ztesting_syrk.c:225-235)
  TRSM  <= 10      (testing/ztesting_trsm.c:246-256)
  POTRF ||A - L L^T|| / (N ||A|| eps) <= 10 and agreement with the oracle's
        tiled Cholesky.
Small-integer inputs make every product and partial sum exact in FP64, so
those cases must match bit for bit.
"""
import numpy as np
import pytest

from helpers import rng, uniform, small_int, spd, tri_matrix

pytestmark = pytest.mark.gpu

TR = [111, 112]


@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("tb", TR)
@pytest.mark.parametrize("shape", [(64, 64, 64), (300, 257, 129), (513, 200, 700), (1, 1, 1), (130, 1, 77)])
def test_dgemm_vs_oracle(xk, orc, ta, tb, shape):
    M, N, K = shape
    r = rng(hash((ta, tb, shape)) % 2**32)
    A = uniform(r, M, K) if ta == 111 else uniform(r, K, M)
    B = uniform(r, K, N) if tb == 111 else uniform(r, N, K)
    C0 = uniform(r, M, N)
    alpha, beta = 1.5, -0.5
    Cref = orc.dgemm_tiled(ta, tb, alpha, A, B, beta, C0.copy(order="F"), 128)
    C = C0.copy(order="F")
    xk.lib().xkblas_set_param(128, 0)
    assert xk.gemm(ta, tb, alpha, A, B, beta, C) == 0
    xk.invalidate()
    res = orc.gemm_residual(alpha, A, B, beta, C0, Cref, C)
    assert res <= 3.0, res


@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("tb", TR)
def test_dgemm_small_int_exact(xk, orc, ta, tb):
    M, N, K = 200, 136, 333
    r = rng(7)
    A = small_int(r, M, K) if ta == 111 else small_int(r, K, M)
    B = small_int(r, K, N) if tb == 111 else small_int(r, N, K)
    C0 = small_int(r, M, N)
    Cref = orc.dgemm_tiled(ta, tb, 2.0, A, B, -1.0, C0.copy(order="F"), 64)
    C = C0.copy(order="F")
    xk.lib().xkblas_set_param(64, 0)
    xk.gemm(ta, tb, 2.0, A, B, -1.0, C)
    xk.invalidate()
    assert np.array_equal(C, Cref)


def test_dgemm_beta_zero_ignores_nan(xk):
    r = rng(3)
    A, B = uniform(r, 100, 50), uniform(r, 50, 70)
    C = np.full((100, 70), np.nan, order="F")
    xk.gemm(111, 111, 1.0, A, B, 0.0, C)
    xk.invalidate()
    assert np.all(np.isfinite(C))
    np.testing.assert_allclose(C, A @ B, rtol=0, atol=1e-13)


def test_dgemm_alpha_zero_scales(xk):
    r = rng(4)
    A, B, C0 = uniform(r, 90, 40), uniform(r, 40, 33), uniform(r, 90, 33)
    C = C0.copy(order="F")
    xk.gemm(111, 111, 0.0, A, B, 3.0, C)
    xk.invalidate()
    np.testing.assert_array_equal(C, 3.0 * C0)


def test_dgemm_config1_shape(xk, orc):
    """The CPU-runnable benchmark case: 2048^3, alpha 1.5, beta -0.5, tile 512."""
    r = rng(1)
    n = 2048
    A, B, C0 = uniform(r, n, n), uniform(r, n, n), uniform(r, n, n)
    path = orc.hostblas()
    if path is None:
        pytest.skip("no host BLAS")
    Cref = orc.host_dgemm(111, 111, 1.5, A, B, -0.5, C0.copy(order="F"))
    C = C0.copy(order="F")
    xk.lib().xkblas_set_param(512, 0)
    xk.gemm(111, 111, 1.5, A, B, -0.5, C)
    xk.invalidate()
    assert orc.gemm_residual(1.5, A, B, -0.5, C0, Cref, C) <= 3.0


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", TR)
@pytest.mark.parametrize("NK", [(300, 200), (64, 500), (257, 1)])
def test_dsyrk_vs_oracle(xk, orc, uplo, trans, NK):
    N, K = NK
    r = rng(11)
    A = uniform(r, N, K) if trans == 111 else uniform(r, K, N)
    C0 = orc.plgsy(0.0, N, 51)  # the reference test's C (testing/ztesting_syrk.c:126)
    Cref = orc.dsyrk_tiled(uplo, trans, 0.75, A, 0.25, C0.copy(order="F"), 128)
    C = C0.copy(order="F")
    xk.lib().xkblas_set_param(128, 0)
    xk.syrk(uplo, trans, 0.75, A, 0.25, C)
    xk.invalidate()
    assert orc.syrk_residual(A, C0, Cref, C, uplo) <= 10.0
    # the other triangle is untouched
    other = np.triu(np.ones((N, N)), 1) if uplo == 122 else np.tril(np.ones((N, N)), -1)
    assert np.array_equal(C[other > 0], C0[other > 0])


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", TR)
@pytest.mark.parametrize("diag", [131, 132])
def test_dtrsm_vs_oracle(xk, orc, side, uplo, ta, diag):
    M, N = 300, 170
    na = M if side == 141 else N
    r = rng(side + uplo + ta + diag)
    A = tri_matrix(r, na, uplo == 122, diag == 132)
    B0 = uniform(r, M, N, 0.0, 1.0)
    alpha = 0.8
    Xref = orc.dtrsm_tiled(side, uplo, ta, diag, alpha, A, B0.copy(order="F"), 128)
    X = B0.copy(order="F")
    xk.lib().xkblas_set_param(128, 0)
    xk.trsm(side, uplo, ta, diag, alpha, A, X)
    xk.invalidate()
    assert orc.trsm_residual(A, uplo, diag, Xref, X) <= 10.0


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("ta", TR)
def test_dtrsm_batched_block_inverses(xk, orc, uplo, ta):
    """Order 1500 with 512-blocks: two full diagonal blocks inverted in one
    batched recursion (three levels) plus a ragged 476 block."""
    M, N = 1500, 64
    r = rng(uplo * 3 + ta)
    A = tri_matrix(r, M, uplo == 122, False)
    B0 = uniform(r, M, N, 0.0, 1.0)
    Xref = orc.dtrsm_tiled(141, uplo, ta, 131, 1.0, A, B0.copy(order="F"), 512)
    X = B0.copy(order="F")
    xk.lib().xkblas_set_param(512, 0)
    xk.trsm(141, uplo, ta, 131, 1.0, A, X)
    xk.invalidate()
    assert orc.trsm_residual(A, uplo, 131, Xref, X) <= 10.0


def test_dtrsm_exact_unit_lower(xk):
    """L unit lower with small integers, X integers: B = L X exactly, solve returns X."""
    r = rng(21)
    n, nrhs = 200, 64
    L = np.tril(r.integers(-1, 2, size=(n, n)).astype(np.float64), -1)
    # keep |entries| of L^-1 and partial sums small: sparse band
    L[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > 1] = 0.0
    np.fill_diagonal(L, 1.0)
    L = np.asfortranarray(L)
    X0 = np.asfortranarray(r.integers(-4, 5, size=(n, nrhs)).astype(np.float64))
    B = np.asfortranarray(L @ X0)
    xk.lib().xkblas_set_param(64, 0)
    xk.trsm(141, 122, 111, 132, 1.0, L, B)
    xk.invalidate()
    # inverse-based blocks: exact up to rounding of the inverse, which is
    # integer here as well
    assert np.array_equal(B, X0)


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("n_nb", [(64, 64), (300, 128), (1000, 256), (777, 512)])
def test_dpotrf_vs_oracle(xk, orc, uplo, n_nb):
    n, nb = n_nb
    r = rng(n)
    A0 = spd(r, n)
    Lref = A0.copy(order="F")
    assert orc.dpotrf_tiled(uplo, Lref, nb) == 0
    A = A0.copy(order="F")
    xk.lib().xkblas_set_param(nb, 0)
    info = xk.potrf(uplo, A)
    xk.invalidate()
    assert info == 0
    assert orc.potrf_residual(A0, A, uplo) <= 10.0
    tri = np.tril if uplo == 122 else np.triu
    np.testing.assert_allclose(tri(A), tri(Lref), rtol=0, atol=1e-12 * np.abs(Lref).max())
    # the other triangle is left as it was (LAPACK semantics)
    other = np.triu(np.ones((n, n)), 1) if uplo == 122 else np.tril(np.ones((n, n)), -1)
    assert np.array_equal(A[other > 0], A0[other > 0])


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("pairing", [0, 1])
@pytest.mark.parametrize("n_nb", [(1000, 128), (1100, 256), (640, 128), (513, 64)])
def test_dpotrf_paired_updates(xk, orc, uplo, pairing, n_nb):
    """Paired trailing updates (two panels per rank-2NB update) and the
    per-panel schedule, forced on and off in FP64, against the oracle: even
    and odd block counts, ragged last blocks, the info of a failing order."""
    n, nb = n_nb
    L = xk.lib()
    r = rng(n + nb)
    A0 = spd(r, n)
    Lref = A0.copy(order="F")
    assert orc.dpotrf_tiled(uplo, Lref, nb) == 0
    A = A0.copy(order="F")
    L.xkblas_set_param(nb, 0)
    L.xkblas_set_potrf_pairing(pairing)
    try:
        info = xk.potrf(uplo, A)
        xk.invalidate()
        assert info == 0
        assert orc.potrf_residual(A0, A, uplo) <= 10.0
        tri = np.tril if uplo == 122 else np.triu
        np.testing.assert_allclose(tri(A), tri(Lref), rtol=0, atol=1e-12 * np.abs(Lref).max())
        other = np.triu(np.ones((n, n)), 1) if uplo == 122 else np.tril(np.ones((n, n)), -1)
        assert np.array_equal(A[other > 0], A0[other > 0])
        bad = n - nb // 2  # inside the last block: reached through every deferred update
        A = A0.copy(order="F")
        A[bad, bad] = -1e6
        info = xk.potrf(uplo, A)
        xk.invalidate()
        assert info == bad + 1
    finally:
        L.xkblas_set_potrf_pairing(-1)


def test_dpotrf_exact_integer_factor(xk):
    """A = L0 L0^T with L0 integer and a power-of-two diagonal: the factor is exact."""
    r = rng(5)
    n = 256
    L0 = np.tril(r.integers(-2, 3, size=(n, n)).astype(np.float64), -1)
    L0[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > 3] = 0.0
    np.fill_diagonal(L0, 8.0)
    A = np.asfortranarray(L0 @ L0.T)
    xk.lib().xkblas_set_param(128, 0)
    info = xk.potrf(122, A)
    xk.invalidate()
    assert info == 0
    np.testing.assert_allclose(np.tril(A), L0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("bad", [5, 130, 299])
def test_dpotrf_info_not_spd(xk, bad):
    r = rng(9)
    n = 300
    A = spd(r, n)
    A[bad, bad] = -1e6  # leading minor of order bad+1 is not positive
    xk.lib().xkblas_set_param(128, 0)
    info = xk.potrf(122, A)
    xk.invalidate()
    assert info == bad + 1


def test_sgemm_vs_fp64(xk):
    r = rng(31)
    M, N, K = 257, 300, 190
    A, B, C0 = (uniform(r, M, K, dtype=np.float32), uniform(r, K, N, dtype=np.float32),
                uniform(r, M, N, dtype=np.float32))
    C = C0.copy(order="F")
    xk.gemm(111, 112 if False else 111, 1.25, A, B, 0.5, C)
    xk.invalidate()
    ref = 1.25 * A.astype(np.float64) @ B.astype(np.float64) + 0.5 * C0.astype(np.float64)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def test_tile_cache_resident_reuse_and_coherence(xk):
    """Second call reuses device tiles; host copy is stale until coherent+sync."""
    L = xk.lib()
    r = rng(41)
    A, B = uniform(r, 256, 256), uniform(r, 256, 256)
    C = np.zeros((256, 256), order="F")
    L.xkblas_set_param(128, 0)
    xk.gemm(111, 111, 1.0, A, B, 0.0, C, sync=False)
    L.xkblas_sync()
    assert np.all(C == 0.0)  # not coherent yet
    xk.coherent(C)
    np.testing.assert_allclose(C, A @ B, atol=1e-12)
    # C <- A B + C again, from device-resident tiles
    xk.gemm(111, 111, 1.0, A, B, 1.0, C)
    np.testing.assert_allclose(C, 2 * (A @ B), atol=1e-12)
    xk.invalidate()


def test_tile_cache_eviction(xk):
    """A small cache limit forces LRU eviction with writeback of dirty tiles."""
    L = xk.lib()
    r = rng(43)
    n = 512
    L.xkblas_set_param(256, 0)
    ev0 = L.xkblas_cache_evictions()
    L.xkblas_set_cache_limit(4 * n * n * 8)  # room for 4 matrices
    mats = [uniform(r, n, n) for _ in range(3)]
    outs = [np.zeros((n, n), order="F") for _ in range(3)]
    for X, Y in zip(mats, outs):
        xk.gemm(111, 111, 1.0, X, X, 0.0, Y, sync=False)
    for Y in outs:
        xk.coherent(Y)
    for X, Y in zip(mats, outs):
        np.testing.assert_allclose(Y, X @ X, atol=1e-11)
    assert L.xkblas_cache_evictions() > ev0
    L.xkblas_set_cache_limit(1 << 40)
    xk.invalidate()


def test_tile_cache_evicts_clean_mirrors_first(xk):
    """With a dirty output less recently used than a clean operand, eviction
    drops the clean one (no write-back, no device-to-host bytes) and the
    results stay right."""
    import ctypes
    L = xk.lib()
    r = rng(44)
    n = 512
    L.xkblas_set_param(256, 0)
    xk.invalidate()  # no mirrors of earlier tests; their scratch stays counted
    L.xkblas_set_cache_limit(L.xkblas_cache_bytes() + 4 * n * n * 8)  # room for 4 matrices
    A, B = uniform(r, n, n), uniform(r, n, n)
    Y, Z, W = (np.zeros((n, n), order="F") for _ in range(3))
    h, d0, d1 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    try:
        xk.gemm(111, 111, 1.0, A, A, 0.0, Y, sync=False)  # A clean, Y dirty
        xk.gemm(111, 111, 1.0, B, B, 0.0, Z, sync=False)  # B clean, Z dirty: the cache is full
        L.xkblas_sync()
        L.xkblas_cache_traffic(ctypes.byref(h), ctypes.byref(d0))
        ev0 = L.xkblas_cache_evictions()
        xk.gemm(111, 111, 1.0, A, A, 0.0, W, sync=False)  # W needs room: B (clean) goes, not Y
        L.xkblas_sync()
        L.xkblas_cache_traffic(ctypes.byref(h), ctypes.byref(d1))
        assert L.xkblas_cache_evictions() > ev0
        assert d1.value == d0.value  # no write-back: the dirty Y and Z stayed
        for X, out in ((A, Y), (B, Z), (A, W)):
            xk.coherent(out)
            np.testing.assert_allclose(out, X @ X, atol=1e-11)
    finally:
        L.xkblas_set_cache_limit(1 << 40)
        xk.invalidate()


def _pinned(L, shape):
    """A Fortran-ordered float64 array in xkblas_malloc (pinned) memory."""
    import ctypes
    nbytes = int(np.prod(shape)) * 8
    p = L.xkblas_malloc(nbytes)
    assert p, "xkblas_malloc failed"
    flat = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(int(np.prod(shape)),))
    return p, nbytes, flat.reshape(shape, order="F")


def test_tile_cache_staging_chunks_and_pinned(xk):
    """Tiles larger than one 64 MiB staging buffer (4096 x 4096 FP64 = 2 chunks
    each, wrapping the 4-buffer ring) round-trip through pageable host memory;
    the same product on xkblas_malloc memory (direct pinned copies) is bit
    identical."""
    L = xk.lib()
    r = rng(47)
    n = 4100
    A, B, C0 = uniform(r, n, n), uniform(r, n, n), uniform(r, n, n)
    nb0 = L.xkblas_get_param()
    L.xkblas_set_param(4096, 0)
    try:
        C = C0.copy(order="F")
        xk.gemm(111, 111, 1.0, A, B, 0.5, C)
        xk.invalidate()
        ref = A @ B + 0.5 * C0
        eps = np.finfo(np.float64).eps
        err = np.abs(C - ref).max() / ((np.abs(A).max() * np.abs(B).max() * n + np.abs(C0).max()) * eps)
        assert err < 3.0, err
        bufs = []
        try:
            for X in (A, B, C0):
                p, nb, Y = _pinned(L, X.shape)
                Y[...] = X
                bufs.append((p, nb, Y))
            Ap, Bp, Cp = (b[2] for b in bufs)
            xk.gemm(111, 111, 1.0, Ap, Bp, 0.5, Cp)
            xk.invalidate()
            assert np.array_equal(Cp, C)
            # pinned operands, pageable output: the write-back goes through the ring
            C2 = C0.copy(order="F")
            xk.gemm(111, 111, 1.0, Ap, Bp, 0.5, C2)
            xk.invalidate()
            assert np.array_equal(C2, C)
        finally:
            for p, nb, _ in bufs:
                L.xkblas_free(p, nb)
    finally:
        L.xkblas_set_param(nb0, 0)
        xk.invalidate()
