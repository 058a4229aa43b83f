/*
 * xkblas_amd_ext.h — extensions of libxkblas_amd.so with no counterpart in
 * the reference API.  Included by xkblas.h.
 */
#ifndef XKBLAS_AMD_EXT_H
#define XKBLAS_AMD_EXT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-launch HIP-event timing of every BLAS call on the compute stream
 * (the role of the reference's XKBLAS_VERBOSE per-format time,
 * blas/xkblas.c:1641-1725).  Classes: "gemm_f64", "syrk_f64", "trsm_f64",
 * "potrf_f64" and the f32 twins, "split_gemm", ...  Values are collected at
 * xkblas_sync(). */
void xkblas_profile_enable(int on);
void xkblas_profile_reset(void);
int  xkblas_profile_get(const char* cls, long* launches, double* total_ms, double* flops);

/* Factor the next diagonal block of POTRF on a side stream while the
 * trailing update runs (default on). */
void xkblas_set_lookahead(int on);

/* Tile-cache introspection and capacity (bytes of HBM; the reference's
 * XKBLAS_CACHE_LIMIT, kaapi/kaapi_impl.c:311-317). */
long xkblas_cache_bytes(void);
long xkblas_cache_evictions(void);
void xkblas_set_cache_limit(size_t bytes);

#ifdef __cplusplus
}
#endif

#endif
