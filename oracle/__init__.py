"""CPU oracle of the tiled BLAS-3 path (test infrastructure only).

Loaded by tests/, the smoke check and the CPU-baseline leg of bench.py as
the checker; the product library never uses it.  See oracle.c for what each
function restates and how it is pinned.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        i, d, p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.oracle_dgemm_tiled.argtypes = [i, i, i, i, i, d, p, i, p, i, d, p, i, i]
        L.oracle_dsyrk_tiled.argtypes = [i, i, i, i, d, p, i, d, p, i, i]
        L.oracle_dtrsm_tiled.argtypes = [i, i, i, i, i, i, d, p, i, p, i, i]
        L.oracle_dpotrf_tiled.argtypes = [i, i, p, i, i]
        L.oracle_tile_dgemm.argtypes = [i, i, i, i, i, d, p, i, p, i, d, p, i]
        L.oracle_tile_dsyrk.argtypes = [i, i, i, i, d, p, i, d, p, i]
        L.oracle_tile_dtrsm.argtypes = [i, i, i, i, i, i, d, p, i, p, i]
        L.oracle_tile_dpotrf.argtypes = [i, i, p, i]
        L.oracle_dgemm_dd.argtypes = [i, i, i, i, i, d, p, i, p, i, d, p, i, p, p]
        L.oracle_plgsy.argtypes = [d, i, p, i, ctypes.c_ulonglong]
        L.oracle_dd_entries.argtypes = [i, i, i, d, p, i, p, i, d, p, i, i, p, p, p, p]
        L.oracle_hostblas_open.argtypes = [ctypes.c_char_p]
        L.oracle_host_dgemm.argtypes = [i, i, i, i, i, d, p, i, p, i, d, p, i]
        L.oracle_host_dsyrk.argtypes = [i, i, i, i, d, p, i, d, p, i]
        L.oracle_host_dtrsm.argtypes = [i, i, i, i, i, i, d, p, i, p, i]
        L.oracle_host_dpotrf.argtypes = [i, i, p, i]
        for f in ("oracle_flops_gemm",):
            getattr(L, f).argtypes = [d, d, d]
            getattr(L, f).restype = d
        L.oracle_flops_syrk.argtypes = [d, d]
        L.oracle_flops_syrk.restype = d
        L.oracle_flops_trsm.argtypes = [i, d, d]
        L.oracle_flops_trsm.restype = d
        L.oracle_flops_trsm_ref.argtypes = [i, d, d]
        L.oracle_flops_trsm_ref.restype = d
        L.oracle_flops_potrf.argtypes = [d]
        L.oracle_flops_potrf.restype = d
        _lib = L
    return _lib


def _p(a):
    assert a.flags.f_contiguous and a.dtype == np.float64
    return a.ctypes.data_as(ctypes.c_void_p)


def dgemm_tiled(ta, tb, alpha, A, B, beta, C, nb):
    M, N = C.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    lib().oracle_dgemm_tiled(ta, tb, M, N, K, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(C),
                             C.shape[0], nb)
    return C


def dsyrk_tiled(uplo, trans, alpha, A, beta, C, nb):
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    lib().oracle_dsyrk_tiled(uplo, trans, N, K, alpha, _p(A), A.shape[0], beta, _p(C), C.shape[0], nb)
    return C


def dtrsm_tiled(side, uplo, ta, diag, alpha, A, B, nb):
    M, N = B.shape
    lib().oracle_dtrsm_tiled(side, uplo, ta, diag, M, N, alpha, _p(A), A.shape[0], _p(B), B.shape[0], nb)
    return B


def dpotrf_tiled(uplo, A, nb):
    return lib().oracle_dpotrf_tiled(uplo, A.shape[0], _p(A), A.shape[0], nb)


def dgemm_dd(ta, tb, alpha, A, B, beta, C):
    M, N = C.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    hi = np.zeros((M, N), order="F")
    lo = np.zeros((M, N), order="F")
    lib().oracle_dgemm_dd(ta, tb, M, N, K, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(C), C.shape[0],
                          _p(hi), _p(lo))
    return hi, lo


def dd_entries(ta, tb, alpha, A, B, beta, C0, rows, cols):
    """Double-double (hi, lo) of entries (rows[e], cols[e]) of alpha op(A) op(B) + beta C0."""
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    hi = np.zeros(len(rows))
    lo = np.zeros(len(rows))
    lib().oracle_dd_entries(ta, tb, K, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(C0), C0.shape[0],
                            len(rows), rows.ctypes.data_as(ctypes.c_void_p), cols.ctypes.data_as(ctypes.c_void_p),
                            hi.ctypes.data_as(ctypes.c_void_p), lo.ctypes.data_as(ctypes.c_void_p))
    return hi, lo


def plgsy(bump, N, seed):
    A = np.zeros((N, N), order="F")
    lib().oracle_plgsy(bump, N, _p(A), N, seed)
    return A


HOSTBLAS_CANDIDATES = [os.environ.get("XKBLAS_BLASLIB", ""), "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so",
                       "libmkl_rt.so.2"]


def hostblas():
    """Open the vendor host BLAS (the reference's native path); returns its path or None."""
    for c in HOSTBLAS_CANDIDATES:
        if c and lib().oracle_hostblas_open(c.encode()) == 0:
            return c
    return None


def hostblas_version(path):
    """Vendor and version string of the host BLAS at `path` (MKL's
    MKL_Get_Version_String, OpenBLAS's openblas_get_config), or its file name."""
    try:
        L = ctypes.CDLL(path)
    except OSError:
        return os.path.basename(path)
    f = getattr(L, "MKL_Get_Version_String", None)
    if f is not None:
        buf = ctypes.create_string_buffer(256)
        f.argtypes = [ctypes.c_char_p, ctypes.c_int]
        f.restype = None
        f(buf, 256)
        return buf.value.decode(errors="replace").strip()
    f = getattr(L, "openblas_get_config", None)
    if f is not None:
        f.restype = ctypes.c_char_p
        return "OpenBLAS " + f().decode(errors="replace").strip()
    return os.path.basename(path)


def host_dgemm(ta, tb, alpha, A, B, beta, C):
    M, N = C.shape
    K = A.shape[1] if ta == NoTrans else A.shape[0]
    assert lib().oracle_host_dgemm(ta, tb, M, N, K, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(C),
                                   C.shape[0]) == 0
    return C


def host_dsyrk(uplo, trans, alpha, A, beta, C):
    N = C.shape[0]
    K = A.shape[1] if trans == NoTrans else A.shape[0]
    assert lib().oracle_host_dsyrk(uplo, trans, N, K, alpha, _p(A), A.shape[0], beta, _p(C), C.shape[0]) == 0
    return C


def host_dtrsm(side, uplo, ta, diag, alpha, A, B):
    M, N = B.shape
    assert lib().oracle_host_dtrsm(side, uplo, ta, diag, M, N, alpha, _p(A), A.shape[0], _p(B), B.shape[0]) == 0
    return B


def host_dpotrf(uplo, A):
    return lib().oracle_host_dpotrf(uplo, A.shape[0], _p(A), A.shape[0])


# ---- residual checks of the reference test drivers (normwise, inf-norm)
EPS = np.finfo(np.float64).eps / 2  # LAPACK dlamch('e') = 2^-53


def inf_norm(X):
    return float(np.max(np.sum(np.abs(X), axis=1))) if X.size else 0.0


def gemm_residual(alpha, A, B, beta, C0, Cref, Cout):
    """testing/ztesting_gemm.c:288-306: passes when <= 3 (units of eps)."""
    K = A.shape[1] if A.shape[0] == Cout.shape[0] else A.shape[0]
    R = inf_norm(Cref - Cout)
    norm = np.sqrt(K + 2.0) * abs(alpha) * inf_norm(A) * inf_norm(B) + np.sqrt(2.0) * abs(beta) * inf_norm(C0)
    return R / norm / EPS if norm else R


def syrk_residual(A, C0, Cref, Cout, uplo):
    """testing/ztesting_syrk.c:225-235 on the uplo triangle: passes when <= 10."""
    N = Cout.shape[0]
    tri = np.tril if uplo == Lower else np.triu
    R = inf_norm(tri(Cref - Cout))
    return R / ((inf_norm(A) + inf_norm(C0)) * N * EPS)


def trsm_residual(A, uplo, diag, Xref, Xout):
    """testing/ztesting_trsm.c:246-256: passes when <= 10."""
    M, N = Xout.shape
    T = np.tril(A) if uplo == Lower else np.triu(A)
    if diag == Unit:
        np.fill_diagonal(T, 1.0)
    R = inf_norm(Xref - Xout)
    return R / ((inf_norm(T) + inf_norm(Xref)) * max(M, N) * EPS)


def potrf_residual(A0, L, uplo):
    """||A - L L^T|| / (N ||A|| eps) on the factor (no reference test exists)."""
    N = A0.shape[0]
    if uplo == Lower:
        F = np.tril(L)
        R = A0 - F @ F.T
    else:
        F = np.triu(L)
        R = A0 - F.T @ F
    return inf_norm(R) / (N * inf_norm(A0) * EPS)
