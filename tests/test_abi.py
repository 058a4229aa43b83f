"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b(xkblas_\w+|[a-z]+_)\s*\(", re.M)


WRAPPER_HEADER = "xkblas_blas_f77.h"


def declared(wrapper=False):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if (os.path.basename(h) == WRAPPER_HEADER) != wrapper:
            continue
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_drop_in_surface():
    names = declared()
    for must in ["xkblas_init", "xkblas_finalize", "xkblas_sync", "xkblas_dgemm_async", "xkblas_dsyrk_async",
                 "xkblas_dtrsm_async", "xkblas_dpotrf_async", "xkblas_memory_coherent_async",
                 "xkblas_memory_invalidate_caches", "xkblas_dgemm_native_"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import xkblas_amd
    lib = xkblas_amd.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_wrapper_exports_the_f77_drop_in():
    """libxkblas_amd_wrapper.so exports dgemm_ ... without touching the GPU
    when loaded (the device is initialised on the first GPU-bound call)."""
    import xkblas_amd
    lib = ctypes.CDLL(os.path.join(ROOT, "xkblas_amd", "libxkblas_amd_wrapper.so"))
    names = declared(wrapper=True)
    assert set(names) >= {"dgemm_", "dsyrk_", "dtrsm_", "dpotrf_"}
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    import xkblas_amd
    missing = [n for n in declared() if n not in xkblas_amd.SIGNATURES]
    assert not missing, missing


def test_query_calls_without_device():
    """Calls that must work before init and without a GPU."""
    import xkblas_amd
    L = xkblas_amd.lib()
    assert L.xkblas_get_ngpus() == 1
    assert L.xkblas_set_ngpus(2) != 0
    assert b"gfx950" in L.xkblas_get_configuration()
    L.xkblas_set_param(512, 0)
    assert L.xkblas_get_param() == 512
    assert L.xkblas_auto_tilesize(None, 1, 2048, 2048, 2048) == 512
    L.xkblas_set_param(0, 0)
    # the reference GEMM heuristic for one GPU (blas/xkblas.c:2502-2515)
    assert L.xkblas_auto_tilesize(None, 1, 16384, 16384, 16384) == 8192
    assert L.xkblas_auto_tilesize(None, 1, 2048, 2048, 2048) == 2048


# (kernel id, M, expected NB), each worked by hand from the reference's
# one-GPU rule (blas/xkblas.c:2437-2521): kernel ids of xkblas_kernel_t.
TILE_CASES = [
    # GEMM (1), GEMMT (2): M, M/2 above 4096, at least 512, 64-aligned
    (1, 300, 512), (1, 4096, 4096), (1, 5000, 2560), (2, 16384, 8192),
    # TRMM (3), TRSM (4): M/2 64-aligned, at least 2048
    (3, 3000, 2048), (4, 10000, 5056),
    # SYMM (5), HEMM (8): M down to a multiple of 1024, at least 1024,
    # doubled below 4096
    (5, 3000, 4096), (8, 16384, 16384), (5, 500, 2048),
    # SYRK (6), HERK (9): M/2 masked to 2048 multiples; M/2 <= 1024 -> M
    # (masked too); above 4096 -> M/4
    (6, 2048, 2048), (6, 2049, 2048), (9, 1000, 1000), (6, 8192, 4096), (6, 20000, 5000), (6, 700, 700),
    # SYR2K (7), HER2K (10): M/4, M/3, M/2 (64-aligned) until above 2048
    (7, 16384, 4096), (7, 8192, 2752), (10, 4096, 2048), (7, 1000, 512), (7, 40000, 5000),
    # default branch (SWAP 11, COPYSCALE 12, POTRF 13): M, M/2 from 4096,
    # at least 1024, 64-aligned
    (11, 3000, 3008), (13, 16384, 8192), (12, 500, 1024),
]


@pytest.mark.parametrize("kernel,m,nb", TILE_CASES)
def test_auto_tilesize_rule(kernel, m, nb):
    import xkblas_amd
    L = xkblas_amd.lib()
    L.xkblas_set_param(0, 0)
    assert L.xkblas_auto_tilesize(None, kernel, m, m // 2, 64) == nb


def test_device_properties_and_registration_without_device():
    import ctypes
    import xkblas_amd
    L = xkblas_amd.lib()
    assert L.xkblas_get_device_properties(None) == 22
    assert L.xkblas_force_set_unified() == 22
    h1 = L.xkblas_register_memory_async(None, 0)
    h2 = L.xkblas_unregister_memory_async(None, 0)
    assert h1 and h2 and h1 != h2
    assert L.xkblas_register_memory_test(h1) == 1
    assert L.xkblas_register_memory_wait(h1) == 0
    assert L.xkblas_register_memory_waitall() == 0
    buf = (ctypes.c_char * 64)()
    assert L.xkblas_register_memory(buf, 64) == 0
    assert L.xkblas_unregister_memory(buf, 64) == 0


def test_split_plans():
    """Moduli and residue-GEMM counts of the split modes (host arithmetic).
    The 2-norm budget per side is T = (log2 P - 1) / 2 for P the product of
    the moduli (||A'_i|| ||B'_j|| <= 2^(2T) < P / 2 bounds every entry of
    the integer product); a row of k equal-magnitude entries keeps
    t = floor(T - log2(k) / 2) bits of its largest entry, and the default
    count is the smallest with t >= 46.  All recomputed here from the sets."""
    import ctypes
    import math
    import xkblas_amd
    L = xkblas_amd.lib()
    bf16 = [511, 509, 507, 505, 503, 499, 493, 491, 487, 479, 473, 467, 463]
    fp8 = [32, 31, 29, 27, 25, 23, 19, 17, 13, 11, 7, 251, 241, 239, 233, 229, 227, 223, 211, 199]
    for mods in (bf16, fp8):  # pairwise coprime, product below 2^123
        assert all(math.gcd(a, b) == 1 for i, a in enumerate(mods) for b in mods[i + 1:])
        assert math.log2(math.prod(mods)) < 123
    assert all(m % 2 == 1 for m in bf16) and max(bf16) <= 511
    assert all(m <= 33 for m in fp8[:11]) and all(m <= 255 for m in fp8)
    nm, npl = ctypes.c_int(), ctypes.c_int()
    # residues within [-255, 255]: exact BF16 values, and 256 products (the
    # run between two folds) plus a remainder stay below 2^24
    assert 256 * 255**2 + 255 < 2**24
    for mode, mods, direct in ((2, bf16, 13), (3, fp8, 11)):
        for k in (64, 1000, 16384, 32768):
            t = L.xkblas_split_info(mode, k, 0, ctypes.byref(nm), ctypes.byref(npl))
            n = nm.value
            assert npl.value == min(n, direct) + 3 * max(0, n - direct)
            T = lambda c: (math.log2(math.prod(mods[:c])) - 1) / 2
            want = lambda c: min(52, math.floor(T(c) - 1e-9 - math.log2(k) / 2))
            assert t == want(n) >= 46
            assert n == 8 or want(n - 1) < 46  # smallest count reaching 46 bits
    assert L.xkblas_split_info(3, 16384, 0, ctypes.byref(nm), ctypes.byref(npl)) == 47
    assert (nm.value, npl.value) == (19, 35)
    assert L.xkblas_split_info(2, 16384, 0, ctypes.byref(nm), ctypes.byref(npl)) == 46
    assert (nm.value, npl.value) == (12, 12)
    assert L.xkblas_split_info(3, 512, 12, ctypes.byref(nm), ctypes.byref(npl)) >= 8
    assert (nm.value, npl.value) == (12, 14)
    assert L.xkblas_split_info(0, 512, 0, None, None) == -1


def test_illegal_argument_returns_code_when_asked():
    """With XKBLAS_ERROR_RETURN=1 the reference's negative codes come back
    (blas/zgemm.c:131-180); checks run before any device use."""
    code = r'''
import ctypes, os, sys
sys.path.insert(0, %r)
import xkblas_amd as x
L = x.lib()
one = (ctypes.c_double * 1)(1.0)
buf = (ctypes.c_double * 16)()
print(L.xkblas_dgemm_async(999, 111, 2, 2, 2, one, buf, 2, buf, 2, one, buf, 2),
      L.xkblas_dgemm_async(111, 111, 4, 2, 2, one, buf, 2, buf, 2, one, buf, 4),
      L.xkblas_dgemm_async(111, 111, 2, 2, 2, one, buf, 2, buf, 2, one, buf, 1),
      L.xkblas_dtrsm_async(141, 999, 111, 131, 2, 2, one, buf, 2, buf, 2),
      L.xkblas_dsyrk_async(122, 111, -1, 2, one, buf, 2, one, buf, 2),
      L.xkblas_dpotrf_async(122, 4, buf, 2))
''' % ROOT
    env = dict(os.environ, XKBLAS_ERROR_RETURN="1")
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["-1", "-8", "-13", "-2", "-3", "-8"]


def test_posv_ir_illegal_arguments_return_codes():
    """xkblas_dposv_ir checks its arguments in LAPACK dsposv's order before
    any device use."""
    code = r'''
import ctypes, sys
sys.path.insert(0, %r)
import xkblas_amd as x
L = x.lib()
a, b, xx = (ctypes.c_double * 16)(), (ctypes.c_double * 16)(), (ctypes.c_double * 16)()
it = ctypes.c_int(7)
print(L.xkblas_dposv_ir(999, 4, 1, a, 4, b, 4, xx, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, -1, 1, a, 4, b, 4, xx, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 4, -1, a, 4, b, 4, xx, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 4, 1, a, 2, b, 4, xx, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 4, 1, a, 4, b, 2, xx, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 4, 1, a, 4, b, 4, xx, 2, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 4, 1, a, 4, b, 4, b, 4, ctypes.byref(it)),
      L.xkblas_dposv_ir(122, 0, 1, a, 1, b, 1, xx, 1, ctypes.byref(it)), it.value)
''' % ROOT
    env = dict(os.environ, XKBLAS_ERROR_RETURN="1")
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["-1", "-2", "-3", "-5", "-7", "-9", "-8", "0", "0"]


def test_more_l3_illegal_arguments_return_codes():
    """GEMMT/SYR2K/SYMM/TRMM argument codes of blas/zgemmt.c:123-159,
    zsyr2k.c:119-150, zsymm.c:115-142, ztrmm.c:116-149."""
    code = r'''
import ctypes, sys
sys.path.insert(0, %r)
import xkblas_amd as x
L = x.lib()
one = (ctypes.c_double * 1)(1.0)
buf = (ctypes.c_double * 64)()
print(L.xkblas_dgemmt_async(122, 999, 111, 2, 2, one, buf, 2, buf, 2, one, buf, 2),
      L.xkblas_dgemmt_async(122, 111, 111, 4, 2, one, buf, 2, buf, 2, one, buf, 4),
      L.xkblas_dgemmt_async(0, 111, 111, 2, 2, one, buf, 2, buf, 2, one, buf, 1),
      L.xkblas_dsyr2k_async(122, 113, 2, 2, one, buf, 2, buf, 2, one, buf, 2),
      L.xkblas_dsyr2k_async(122, 111, 4, 2, one, buf, 4, buf, 2, one, buf, 4),
      L.xkblas_dsymm_async(141, 122, 4, 2, one, buf, 2, buf, 4, one, buf, 4),
      L.xkblas_dsymm_async(141, 122, 2, 2, one, buf, 2, buf, 2, one, buf, 1),
      L.xkblas_dtrmm_async(141, 122, 111, 999, 2, 2, one, buf, 2, buf, 2),
      L.xkblas_dtrmm_async(142, 122, 111, 131, 2, 4, one, buf, 2, buf, 2))
''' % ROOT
    env = dict(os.environ, XKBLAS_ERROR_RETURN="1")
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["-1", "-8", "-13", "-2", "-9", "-7", "-12", "-4", "-8"]


def test_illegal_argument_aborts_by_default():
    """kaapi_error semantics: print and abort (blas/common.gen:312-316)."""
    code = r'''
import ctypes, sys
sys.path.insert(0, %r)
import xkblas_amd as x
L = x.lib()
one = (ctypes.c_double * 1)(1.0)
buf = (ctypes.c_double * 16)()
L.xkblas_dgemm_async(999, 111, 2, 2, 2, one, buf, 2, buf, 2, one, buf, 2)
''' % ROOT
    env = {k: v for k, v in os.environ.items() if k != "XKBLAS_ERROR_RETURN"}
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode != 0
    assert "illegal value of transA" in out.stderr


def test_native_forwarders_of_gemmt_syr2k_symm_trmm():
    """xkblas_d{gemmt,syr2k,symm,trmm}_native run the host BLAS (no GPU):
    checked against their textbook definitions."""
    import numpy as np
    import xkblas_amd
    L = xkblas_amd.lib()
    r = np.random.default_rng(5)
    n, k = 90, 40
    A, B = (np.asfortranarray(r.uniform(-1, 1, (n, k))) for _ in range(2))
    C0 = np.asfortranarray(r.uniform(-1, 1, (n, n)))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    one, half = np.array([1.0]), np.array([0.5])
    C = C0.copy(order="F")
    L.xkblas_dgemmt_native(122, 111, 112, n, k, p(one), p(A), n, p(B), n, p(half), p(C), n)
    assert np.allclose(np.tril(C), np.tril(A @ B.T + 0.5 * C0), atol=1e-13)
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))
    C = C0.copy(order="F")
    L.xkblas_dsyr2k_native(121, 111, n, k, p(one), p(A), n, p(B), n, p(half), p(C), n)
    assert np.allclose(np.triu(C), np.triu(A @ B.T + B @ A.T + 0.5 * C0), atol=1e-13)
    S = np.asfortranarray(C0 + C0.T)
    C = np.zeros((n, k), order="F")
    L.xkblas_dsymm_native(141, 122, n, k, p(one), p(np.asfortranarray(np.tril(S))), n, p(A), n, p(half), p(C), n)
    assert np.allclose(C, S @ A, atol=1e-13)
    T = np.asfortranarray(np.triu(C0))
    X = A.copy(order="F")
    L.xkblas_dtrmm_native(141, 121, 112, 131, n, k, p(one), p(T), n, p(X), n)
    assert np.allclose(X, T.T @ A, atol=1e-13)


def test_complex_native_forwarders():
    """xkblas_{z,c}*_native run the host BLAS (no GPU): checked against their
    textbook definitions."""
    import numpy as np
    import xkblas_amd
    L = xkblas_amd.lib()
    r = np.random.default_rng(7)
    n, k = 60, 30
    c = lambda m, q, dt=np.complex128: np.asfortranarray((r.uniform(-1, 1, (m, q)) + 1j * r.uniform(-1, 1, (m, q))).astype(dt))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    A, B, C0 = c(n, k), c(k, n), c(n, n)
    al, be = np.array([1 - 2j]), np.array([0.5j])
    C = C0.copy(order="F")
    L.xkblas_zgemm_native(111, 111, n, n, k, p(al), p(A), n, p(B), k, p(be), p(C), n)
    assert np.allclose(C, (1 - 2j) * A @ B + 0.5j * C0, atol=1e-12)
    C = C0.copy(order="F")
    L.xkblas_zherk_native(122, 111, n, k, p(np.array([2.0])), p(A), n, p(np.array([1.0])), p(C), n)
    ref = 2.0 * A @ A.conj().T + C0
    np.fill_diagonal(ref, ref.diagonal().real)
    assert np.allclose(np.tril(C), np.tril(ref), atol=1e-12)
    T = np.asfortranarray(np.triu(c(n, n)) + 4 * np.eye(n))
    X = c(n, k)
    X0 = X.copy()
    L.xkblas_ztrsm_native(141, 121, 113, 131, n, k, p(np.array([1.0 + 0j])), p(T), n, p(X), n)
    assert np.allclose(T.conj().T @ X, X0, atol=1e-12)
    G = c(n, n)
    H = np.asfortranarray(G @ G.conj().T + n * np.eye(n))
    F = H.copy(order="F")
    assert L.xkblas_zpotrf_native(122, n, p(F), n) == 0
    assert np.allclose(np.tril(F) @ np.tril(F).conj().T, H, atol=1e-10)
    A32, B32 = c(n, k, np.complex64), c(k, n, np.complex64)
    C32 = np.zeros((n, n), dtype=np.complex64, order="F")
    L.xkblas_cgemm_native(111, 111, n, n, k, p(np.array([1], np.complex64)), p(A32), n, p(B32), k,
                          p(np.array([0], np.complex64)), p(C32), n)
    assert np.allclose(C32, A32 @ B32, atol=1e-4)


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus N > 1 exits with a message, before any
    device use."""
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
