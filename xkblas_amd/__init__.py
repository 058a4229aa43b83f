"""Python binding of libxkblas_amd.so (the xkblas.h C API) over ctypes.

The product is the C-ABI library; this module only declares its symbols and
offers thin helpers that call them on NumPy arrays (Fortran order), the way
the reference's test drivers call the C API (testing/ztesting_gemm.c:157-162:
async -> coherent_async -> sync).  Loading fails loudly when the library has
not been built: there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxkblas_amd.so")

# CBLAS enum values (reference blas/xkblas.h:114-119)
NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

DEFAULT_MATH, TENSOR_OP_MATH, SPLIT_BF16_MATH, SPLIT_FP8_MATH = 0, 1, 2, 3

# every symbol include/*.h declares, with (restype, argtypes)
_i, _d, _f, _p, _sz, _l = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_long
_u64 = ctypes.c_uint64
SIGNATURES = {
    "xkblas_init": (_i, []),
    "xkblas_finalize": (_i, []),
    "xkblas_sync": (_i, []),
    "xkblas_set_param": (None, [_sz, _sz]),
    "xkblas_get_param": (_sz, []),
    "xkblas_get_ngpus": (_i, []),
    "xkblas_set_ngpus": (_i, [_i]),
    "xkblas_get_devicecount": (_i, []),
    "xkblas_get_device_count": (_i, [_p]),
    "xkblas_set_modemath": (None, [_i]),
    "xkblas_get_configuration": (ctypes.c_char_p, []),
    "xkblas_load_sym": (None, [_p, ctypes.c_char_p]),
    "xkblas_elapsedns": (_u64, []),
    "xkblas_elapsedtime": (_d, []),
    "xkblas_auto_tilesize": (_sz, [_p, _i, _sz, _sz, _sz]),
    "xkblas_malloc": (_p, [_sz]),
    "xkblas_free": (None, [_p, _sz]),
    "xkblas_memory_coherent_async": (_i, [_i, _i, _sz, _sz, _p, _sz, _sz]),
    "xkblas_get_device_properties": (_i, [_p]),
    "xkblas_force_set_unified": (_i, []),
    "xkblas_register_memory_async": (_u64, [_p, _sz]),
    "xkblas_unregister_memory_async": (_u64, [_p, _sz]),
    "xkblas_register_memory_test": (_i, [_u64]),
    "xkblas_register_memory_wait": (_i, [_u64]),
    "xkblas_register_memory_waitall": (_i, []),
    "xkblas_register_memory": (_i, [_p, _sz]),
    "xkblas_unregister_memory": (_i, [_p, _sz]),
    "xkblas_memory_syncall": (_i, []),
    "xkblas_memory_invalidate": (_i, [_p]),
    "xkblas_memory_invalidate_caches": (_i, []),
    "xkblas_memory_free": (_i, []),
    "xkblas_dgemm_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsyrk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_dtrsm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_dpotrf_async": (_i, [_i, _i, _p, _i]),
    "xkblas_dgemm_native": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsyrk_native": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_dtrsm_native": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_dpotrf_native": (_i, [_i, _i, _p, _i]),
    "xkblas_zgemm_native": (_i, [_i]*5 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zgemmt_native": (_i, [_i]*5 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ztrsm_native": (_i, [_i]*6 + [_p, _p, _i, _p, _i]),
    "xkblas_ztrmm_native": (_i, [_i]*6 + [_p, _p, _i, _p, _i]),
    "xkblas_zsymm_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zhemm_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zsyrk_native": (_i, [_i]*4 + [_p, _p, _i, _p, _p, _i]),
    "xkblas_zherk_native": (_i, [_i]*4 + [_p, _p, _i, _p, _p, _i]),
    "xkblas_zsyr2k_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zher2k_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zpotrf_native": (_i, [_i, _i, _p, _i]),
    "xkblas_cgemm_native": (_i, [_i]*5 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cgemmt_native": (_i, [_i]*5 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ctrsm_native": (_i, [_i]*6 + [_p, _p, _i, _p, _i]),
    "xkblas_ctrmm_native": (_i, [_i]*6 + [_p, _p, _i, _p, _i]),
    "xkblas_csymm_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_chemm_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_csyrk_native": (_i, [_i]*4 + [_p, _p, _i, _p, _p, _i]),
    "xkblas_cherk_native": (_i, [_i]*4 + [_p, _p, _i, _p, _p, _i]),
    "xkblas_csyr2k_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cher2k_native": (_i, [_i]*4 + [_p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cpotrf_native": (_i, [_i, _i, _p, _i]),
    "xkblas_dgemmt_native": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsyr2k_native": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsymm_native": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dtrmm_native": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_sgemm_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ssyrk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_strsm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_spotrf_async": (_i, [_i, _i, _p, _i]),
    "xkblas_potrf_info": (_i, [_p]),
    "xkblas_set_split_count": (None, [_i]),
    "xkblas_get_split_count": (_i, []),
    "xkblas_profile_enable": (None, [_i]),
    "xkblas_profile_reset": (None, []),
    "xkblas_trace_start": (_i, [ctypes.c_char_p]),
    "xkblas_trace_stop": (_i, []),
    "xkblas_profile_get": (_i, [ctypes.c_char_p, _p, _p, _p]),
    "xkblas_set_lookahead": (None, [_i]),
    "xkblas_set_pipeline": (None, [_i]),
    "xkblas_cache_bytes": (_l, []),
    "xkblas_cache_evictions": (_l, []),
    "xkblas_cache_traffic": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "xkblas_set_cache_limit": (None, [_sz]),
    "xkblas_batch_create": (_p, [_p, _i]),
    "xkblas_batch_run": (_i, [_p]),
    "xkblas_batch_run_noalloc": (_i, [_p, _i]),
    "xkblas_set_batch_tile": (None, [_i]),
    "xkblas_batch_use_graph": (_i, [_p, _i]),
    "xkblas_batch_info": (_i, [_p, _p, _p, _p, _p]),
    "xkblas_batch_destroy": (None, [_p]),
    "xkblas_device_malloc": (_p, [_sz]),
    "xkblas_device_free": (None, [_p]),
    "xkblas_memcpy_htod": (_i, [_p, _p, _sz]),
    "xkblas_memcpy_dtoh": (_i, [_p, _p, _sz]),
    "xkblas_mlp_create": (_p, [_i, ctypes.c_uint]),
    "xkblas_mlp_num_params": (_l, [_p]),
    "xkblas_mlp_set_params": (_i, [_p, _p]),
    "xkblas_mlp_get_params": (_i, [_p, _p]),
    "xkblas_mlp_forward": (_i, [_p, _p, _i, _p]),
    "xkblas_mlp_gradient": (_i, [_p, _p, _p, _i, _p, _p]),
    "xkblas_mlp_train": (_f, [_p, _p, _p, _i, _i, _i, _f]),
    "xkblas_mlp_destroy": (None, [_p]),
    "xkblas_tune_time": (_d, [_i] * 13),
    "xkblas_tune_release": (None, []),
    "xkblas_set_gemm_tile": (None, [_i]),
    "xkblas_set_split_kernel": (None, [_i]),
    "xkblas_set_potrf_pairing": (None, [_i]),
    "xkblas_zgemm_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cgemm_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zsyrk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_zsyr2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zher2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zgemmt_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_zsymm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ztrmm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_csyrk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_csyr2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cher2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_cgemmt_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_csymm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ctrmm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_zpotrf_async": (_i, [_i, _i, _p, _i]),
    "xkblas_cpotrf_async": (_i, [_i, _i, _p, _i]),
    "xkblas_ztrsm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_ctrsm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_zherk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_cherk_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _p, _i]),
    "xkblas_zhemm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_chemm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dgemmt_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_sgemmt_async": (_i, [_i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsyr2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ssyr2k_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dsymm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_ssymm_async": (_i, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i]),
    "xkblas_dtrmm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_strmm_async": (_i, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i]),
    "xkblas_dposv_ir": (_i, [_i, _i, _i, _p, _i, _p, _i, _p, _i, _p]),
    "xkblas_set_posv_ir": (None, [_i, _i]),
    "xkblas_split_side_bits": (_i, [_i, _i]),
    "xkblas_split_info": (_i, [_i, _i, _i, _p, _p]),
    "xkblas_dgemm_native_": (None, [_p] * 13),
    "xkblas_dsyrk_native_": (None, [_p] * 10),
    "xkblas_dtrsm_native_": (None, [_p] * 11),
    "xkblas_dpotrf_native_": (None, [_p] * 5),
    "xkblas_zgemm_native_": (None, [_p] * 13),
    "xkblas_zgemmt_native_": (None, [_p] * 13),
    "xkblas_ztrsm_native_": (None, [_p] * 11),
    "xkblas_ztrmm_native_": (None, [_p] * 11),
    "xkblas_zsymm_native_": (None, [_p] * 12),
    "xkblas_zhemm_native_": (None, [_p] * 12),
    "xkblas_zsyrk_native_": (None, [_p] * 10),
    "xkblas_zherk_native_": (None, [_p] * 10),
    "xkblas_zsyr2k_native_": (None, [_p] * 12),
    "xkblas_zher2k_native_": (None, [_p] * 12),
    "xkblas_zpotrf_native_": (None, [_p] * 5),
    "xkblas_cgemm_native_": (None, [_p] * 13),
    "xkblas_cgemmt_native_": (None, [_p] * 13),
    "xkblas_ctrsm_native_": (None, [_p] * 11),
    "xkblas_ctrmm_native_": (None, [_p] * 11),
    "xkblas_csymm_native_": (None, [_p] * 12),
    "xkblas_chemm_native_": (None, [_p] * 12),
    "xkblas_csyrk_native_": (None, [_p] * 10),
    "xkblas_cherk_native_": (None, [_p] * 10),
    "xkblas_csyr2k_native_": (None, [_p] * 12),
    "xkblas_cher2k_native_": (None, [_p] * 12),
    "xkblas_cpotrf_native_": (None, [_p] * 5),
    "xkblas_dgemmt_native_": (None, [_p] * 13),
    "xkblas_dsyr2k_native_": (None, [_p] * 12),
    "xkblas_dsymm_native_": (None, [_p] * 12),
    "xkblas_dtrmm_native_": (None, [_p] * 11),
}

_lib = None


def lib():
    """Load libxkblas_amd.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libxkblas_amd.so is not built: run __graft_entry__.build() "
                               "or python -m xkblas_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(a, dtype):
    if not (a.flags.f_contiguous and a.dtype == dtype):
        raise ValueError("expected a Fortran-ordered %s array" % np.dtype(dtype).name)


_PREFIX = {np.dtype(np.float64): "d", np.dtype(np.float32): "s", np.dtype(np.complex128): "z",
           np.dtype(np.complex64): "c"}


def _scalar(v, dtype):
    return np.array([v], dtype=dtype)


def gemm(ta, tb, alpha, A, B, beta, C, sync=True):
    """C <- alpha op(A) op(B) + beta C through xkblas_?gemm_async (+ coherent + sync)."""
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    M, N = C.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    fn = {np.dtype(np.float64): lib().xkblas_dgemm_async, np.dtype(np.float32): lib().xkblas_sgemm_async,
          np.dtype(np.complex128): lib().xkblas_zgemm_async, np.dtype(np.complex64): lib().xkblas_cgemm_async}[dt]
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(ta, tb, M, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, B.shape[0]), _ptr(be), _ptr(C),
            max(1, M))
    if sync:
        coherent(C)
    return rc


def syrk(uplo, trans, alpha, A, beta, C, sync=True):
    dt = C.dtype
    for X in (A, C):
        _check(X, dt)
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    fn = getattr(lib(), "xkblas_%ssyrk_async" % _PREFIX[np.dtype(dt)])
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(uplo, trans, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(be), _ptr(C), max(1, N))
    if sync:
        coherent(C, uplo)
    return rc


def trsm(side, uplo, ta, diag, alpha, A, B, sync=True):
    dt = B.dtype
    for X in (A, B):
        _check(X, dt)
    M, N = B.shape
    fn = getattr(lib(), "xkblas_%strsm_async" % _PREFIX[np.dtype(dt)])
    al = _scalar(alpha, dt)
    rc = fn(side, uplo, ta, diag, M, N, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, M))
    if sync:
        coherent(B)
    return rc


def potrf(uplo, A, sync=True):
    """Returns the LAPACK info (after sync) or the enqueue code."""
    dt = A.dtype
    _check(A, dt)
    N = A.shape[0]
    fn = getattr(lib(), "xkblas_%spotrf_async" % _PREFIX[np.dtype(dt)])
    rc = fn(uplo, N, _ptr(A), max(1, N))
    if rc != 0 or not sync:
        return rc
    coherent(A, uplo)
    return lib().xkblas_potrf_info(_ptr(A))


def gemmt(uplo, ta, tb, alpha, A, B, beta, C, sync=True):
    """C <- alpha op(A) op(B) + beta C on the uplo triangle (0: all of C)."""
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    N = C.shape[0]
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    fn = getattr(lib(), "xkblas_%sgemmt_async" % _PREFIX[np.dtype(dt)])
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(uplo, ta, tb, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, B.shape[0]), _ptr(be), _ptr(C),
            max(1, N))
    if sync:
        coherent(C)
    return rc


def syr2k(uplo, trans, alpha, A, B, beta, C, sync=True):
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    fn = getattr(lib(), "xkblas_%ssyr2k_async" % _PREFIX[np.dtype(dt)])
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(uplo, trans, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, B.shape[0]), _ptr(be), _ptr(C),
            max(1, N))
    if sync:
        coherent(C, uplo)
    return rc


def symm(side, uplo, alpha, A, B, beta, C, sync=True):
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    M, N = C.shape
    fn = getattr(lib(), "xkblas_%ssymm_async" % _PREFIX[np.dtype(dt)])
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(side, uplo, M, N, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, M), _ptr(be), _ptr(C), max(1, M))
    if sync:
        coherent(C)
    return rc


def trmm(side, uplo, ta, diag, alpha, A, B, sync=True):
    dt = B.dtype
    for X in (A, B):
        _check(X, dt)
    M, N = B.shape
    fn = getattr(lib(), "xkblas_%strmm_async" % _PREFIX[np.dtype(dt)])
    al = _scalar(alpha, dt)
    rc = fn(side, uplo, ta, diag, M, N, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, M))
    if sync:
        coherent(B)
    return rc


def herk(uplo, trans, alpha, A, beta, C, sync=True):
    """Complex C <- alpha op(A) op(A)^H + beta C on the uplo triangle (alpha,
    beta real)."""
    dt = C.dtype
    for X in (A, C):
        _check(X, dt)
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    real = np.float64 if dt == np.complex128 else np.float32
    fn = lib().xkblas_zherk_async if dt == np.complex128 else lib().xkblas_cherk_async
    al, be = _scalar(alpha, real), _scalar(beta, real)
    rc = fn(uplo, trans, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(be), _ptr(C), max(1, N))
    if sync:
        coherent(C, uplo)
    return rc


def hemm(side, uplo, alpha, A, B, beta, C, sync=True):
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    M, N = C.shape
    fn = lib().xkblas_zhemm_async if dt == np.complex128 else lib().xkblas_chemm_async
    al, be = _scalar(alpha, dt), _scalar(beta, dt)
    rc = fn(side, uplo, M, N, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, M), _ptr(be), _ptr(C), max(1, M))
    if sync:
        coherent(C)
    return rc


def her2k(uplo, trans, alpha, A, B, beta, C, sync=True):
    """Complex C <- alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C
    (beta real) on the uplo triangle."""
    dt = C.dtype
    for X in (A, B, C):
        _check(X, dt)
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    real = np.float64 if dt == np.complex128 else np.float32
    fn = getattr(lib(), "xkblas_%sher2k_async" % _PREFIX[np.dtype(dt)])
    al, be = _scalar(alpha, dt), _scalar(beta, real)
    rc = fn(uplo, trans, N, K, _ptr(al), _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, B.shape[0]), _ptr(be), _ptr(C),
            max(1, N))
    if sync:
        coherent(C, uplo)
    return rc


def posv_ir(uplo, A, B):
    """X solving A X = B (A SPD, uplo triangle) by xkblas_dposv_ir; returns
    (X, info, iter) with X coherent in host memory."""
    _check(A, np.float64)
    _check(B, np.float64)
    N, NRHS = B.shape
    X = np.zeros((N, NRHS), order="F")
    it = ctypes.c_int(0)
    info = lib().xkblas_dposv_ir(uplo, N, NRHS, _ptr(A), max(1, A.shape[0]), _ptr(B), max(1, N), _ptr(X), max(1, N),
                                 ctypes.byref(it))
    if info >= 0 and N and NRHS:
        coherent(X)
    # X is this call's own array: drop its device mirror, so that a later
    # array at the same address is not taken for it
    lib().xkblas_memory_invalidate(_ptr(X))
    return X, info, it.value


def coherent(X, uplo=0):
    L = lib()
    L.xkblas_memory_coherent_async(uplo, 0, X.shape[0], X.shape[1], _ptr(X), max(1, X.shape[0]), X.itemsize)
    L.xkblas_sync()


def invalidate():
    lib().xkblas_memory_invalidate_caches()


BATCH_GEMM, BATCH_TRSM = 0, 1


class BatchProblem(ctypes.Structure):
    """xkblas_batch_problem_t (include/xkblas_amd_ext.h)."""
    _fields_ = [("kind", ctypes.c_int), ("dtype", ctypes.c_int), ("transA", ctypes.c_int), ("transB", ctypes.c_int),
                ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_long),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_long),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_long)]


class DeviceArray:
    """A Fortran-ordered device copy of a NumPy array (xkblas_device_malloc)."""

    def __init__(self, host):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = lib().xkblas_device_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError("xkblas_device_malloc failed")
        self.upload(host)

    def upload(self, host):
        h = np.asfortranarray(host, dtype=self.dtype)
        if lib().xkblas_memcpy_htod(self.ptr, h.ctypes.data_as(ctypes.c_void_p), self.nbytes) != 0:
            raise RuntimeError("host-to-device copy failed")

    def download(self):
        h = np.empty(self.shape, dtype=self.dtype, order="F")
        if lib().xkblas_memcpy_dtoh(h.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes) != 0:
            raise RuntimeError("device-to-host copy failed")
        return h

    def free(self):
        if self.ptr:
            lib().xkblas_device_free(self.ptr)
            self.ptr = None


def profile(cls):
    n, ms, fl = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
    lib().xkblas_profile_get(cls.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl))
    return n.value, ms.value, fl.value
