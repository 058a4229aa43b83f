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

/* Trace in the CSV schema of the reference's trace reader (bin/ukilli.cpp:
 * 848-1113, the files it writes from a KAAPI_RECORD_TRACE event file):
 * kern.csv and task.csv get one row per call's device work (bracketed by
 * HIP events on the compute stream; Name = the profiling class), cpy.csv one
 * row per host<->device copy (Dir 1 = H2D, 2 = D2H, Size in bytes; Stream 2 =
 * H2D, 3 = D2H), parallel.csv the traced region.  Times are seconds from the
 * start of the trace; rows are written at each xkblas_sync().  Also started at
 * initialisation by XKBLAS_TRACE_DIR=<dir>, and stopped by xkblas_finalize.
 * Returns 0, or -1 if the files cannot be created. */
int  xkblas_trace_start(const char* dir);
int  xkblas_trace_stop(void);

/* Factor the next diagonal block of POTRF on a side stream while the
 * trailing update runs (default on). */
void xkblas_set_lookahead(int on);

/* A DGEMM/SGEMM whose op(A) or C still has tiles in host memory runs by row
 * panels of C (one tile row each): panel p waits only for B and its own rows
 * of op(A) and C, so the host-to-device copies of panel p+1 overlap the
 * product of panel p, and xkblas_memory_coherent_async copies each panel's
 * rows back as soon as they are final.  Same result bits as one launch.
 * Default on; 0 turns it off (environment XKBLAS_PIPELINE). */
void xkblas_set_pipeline(int on);

/* Tile-cache introspection and capacity (bytes of HBM; the reference's
 * XKBLAS_CACHE_LIMIT, kaapi/kaapi_impl.c:311-317). */
long xkblas_cache_bytes(void);
long xkblas_cache_evictions(void);
/* Bytes copied host-to-device and device-to-host by the tile cache since
 * initialisation (no reference counterpart: the reference reports copies in
 * its trace, cpy.csv).  Eviction drops clean mirrors (every tile also valid
 * on the host) before dirty ones, which need a write-back. */
void xkblas_cache_traffic(double* h2d_bytes, double* d2h_bytes);
void xkblas_set_cache_limit(size_t bytes);

/* ---- batched uneven problems in one persistent launch -------------------
 * Problems live in device memory (xkblas_device_malloc or any HIP allocation).
 * GEMM: C <- alpha op(A) op(B) + beta C (op = transA/transB, C m x n, k).
 * TRSM: Left, Lower, NoTrans, Non-unit: L X = alpha B with L = A (m x m,
 *       lda) and B = C (m x n, ldc) overwritten by X.
 * dtype 0 = FP64, 1 = FP32 (mixed in one batch).  A plan holds the work list
 * (64 x 64 or 32 x 32 GEMM tiles and 64-column TRSM panels, longest first)
 * on the device; xkblas_batch_run enqueues one counter fill and one persistent
 * kernel on the library's compute stream (xkblas_sync waits for it);
 * xkblas_batch_use_graph(plan, 1) captures that pair into a hipGraph that
 * later runs replay. */
#define XKBLAS_BATCH_GEMM 0
#define XKBLAS_BATCH_TRSM 1
typedef struct xkblas_batch_problem {
  int kind, dtype, transA, transB, m, n, k;
  double alpha, beta;
  const void* A;
  long lda;
  const void* B;
  long ldb;
  void* C;
  long ldc;
} xkblas_batch_problem_t;

void* xkblas_batch_create(const xkblas_batch_problem_t* problems, int count);
/* GEMM work-item shape of plans created afterwards: 1 = 32 x 32 tiles for a
 * problem whose 32 x 32 tiling covers at most 3/4 of the area of its 64 x 64
 * one (small m or n), 2 = 32 x 32 only, 3 = 64 x 64 only, 0 (default) = rule
 * 1 for FP64 problems and 64 x 64 for FP32 ones; environment
 * XKBLAS_BATCH_TILE.  Results are the same bit for bit. */
void  xkblas_set_batch_tile(int mode);
int   xkblas_batch_run(void* plan);
int   xkblas_batch_use_graph(void* plan, int on);
int   xkblas_batch_info(void* plan, int* nitems, int* grid, double* flops, double* bytes);
void  xkblas_batch_destroy(void* plan);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Runs the problems like xkblas_batch_create +
 * xkblas_batch_run + xkblas_batch_destroy, with the same kernels and results,
 * but allocates nothing: the work counter, problem list and work list go to a
 * fixed device array of 3.9 MB per GPU (held until the process ends) and the
 * host-side lists to fixed host arrays.  At most 8192 problems and 196608
 * work items (64 x 64 GEMM tiles plus 64-column TRSM panels): beyond that it
 * prints the limit and returns -2 (problems) or -3 (items), before touching
 * the GPU.  -1: bad arguments; -4: xkblas_init not called.  Asynchronous like
 * xkblas_batch_run (xkblas_sync waits); the problem array may be reused as
 * soon as the call returns.  Runs share the array: they are ordered on the
 * library's compute stream, so calls from several threads serialise, but work
 * that overlaps a run on the same GPU from another stream would race. */
int   xkblas_batch_run_noalloc(const xkblas_batch_problem_t* problems, int count);

/* ---- tile-cost model: 3-layer MLP (in -> 64 -> 64 -> 1, ReLU, FP32) ------
 * trained on the GPU by the library's own forward / backward / Adam kernels.
 * Parameters are one FP32 vector [W1 (in x 64), b1, W2 (64 x 64), b2, w3, b3];
 * samples are row-major (n x in), in <= 16. */
void* xkblas_mlp_create(int in_features, unsigned seed);
long  xkblas_mlp_num_params(void* mlp);
int   xkblas_mlp_set_params(void* mlp, const float* params);
int   xkblas_mlp_get_params(void* mlp, float* params);
int   xkblas_mlp_forward(void* mlp, const float* X, int n, float* Y);
int   xkblas_mlp_gradient(void* mlp, const float* X, const float* T, int n, float* grad, float* loss);
float xkblas_mlp_train(void* mlp, const float* X, const float* T, int n, int epochs, int batch, float lr);
void  xkblas_mlp_destroy(void* mlp);

/* ---- autotuner timing service and the knobs the cost model sets ---------
 * op: 0 GEMM (C m x n, inner k), 1 SYRK (lower, C n x n, inner k),
 * 2 TRSM (Left, Lower, NoTrans, Non-unit, B m x n), 3 POTRF (lower, order m).
 * Operands are device-resident and filled by a kernel; returns the best
 * kernel time over `iters` runs in ms. */
double xkblas_tune_time(int op, int dtype, int mode, int transA, int transB, int m, int n, int k, int tile, int nb,
                        int split_count, int split_kernel, int iters);
void xkblas_tune_release(void);
void xkblas_set_gemm_tile(int tile);        /* 0 = automatic, 32 (FP64), 64, 128 */
void xkblas_set_split_kernel(int variant);  /* BF16 residue GEMM: 0 = 8 waves of 128x64,
                                             1 = 4 waves of 128x128 */
int  xkblas_split_side_bits(int k, int nsplit);  /* split-BF16 bits per side t */
/* split mode (XKBLAS_SPLIT_BF16_MATH / XKBLAS_SPLIT_FP8_MATH), inner dimension
 * k and requested count (0 = automatic: the smallest count with t >= 46):
 * returns the bits per side t, stores the number of CRT moduli and of residue
 * GEMMs ("planes"); -1 for another mode.  Rows of op(A) and columns of op(B)
 * are scaled by powers of two so that their 2-norms stay within 2^T, 2^(2T) <
 * P / 2 (P the product of the moduli); t = floor(T - log2(k) / 2) is what a
 * row of k equal-magnitude entries keeps of its largest entry (its rounding
 * error is below 2^-t of it); rows with spread magnitudes keep more. */
int  xkblas_split_info(int mode, int k, int nsplit, int* nmod, int* nplane);
/* DPOTRF/SPOTRF trailing updates: 1 = paired (default; -1 selects it too):
 * every other step applies two panels as one rank-2NB update of the trailing
 * triangle, half at once and half at the next step, while the next block
 * column is still updated at once for the lookahead; 0 = one rank-NB update
 * per panel.  Environment XKBLAS_POTRF_PAIR sets the initial value. */
void xkblas_set_potrf_pairing(int mode);

/* Mixed-precision solve of A X = B, A symmetric positive definite (the uplo
 * triangle of A is read), B and X N x NRHS (X must not be B).  No counterpart
 * in the reference; argument order, codes and *iter follow LAPACK dsposv:
 * the Cholesky factor is computed in split-BF16 mode with a reduced number of
 * CRT moduli (diagonal blocks in FP64), the residual B - A X and the updates
 * are FP64.  Refinement stops when every column passes dsposv's test on the
 * residual the last correction d was computed from,
 * ||r||_inf <= sqrt(N) ||x||_inf ||A||_inf eps, and a forward test: ||d||_inf
 * <= 2^-48 ||x||_inf, or (from the second correction) rho / (1 - rho) ||d||_inf
 * <= 2^-48 ||x||_inf with rho = ||d|| / ||previous d||, or rho > 1/2 (no
 * further progress).  dsposv's test alone allows a relative error of about
 * sqrt(N) eps cond(A).  *iter = corrections applied;
 * -3 if the reduced factor failed, -(itermax + 1) after itermax corrections:
 * then X comes from the factor of the current mode (as dsposv falls back to
 * dpotrf).  Returns 0, > 0 (LAPACK info: A not positive definite) or < 0
 * (argument).  Synchronous; X is then valid on the device and reaches host
 * memory by xkblas_memory_coherent_async + xkblas_sync, like every result. */
int  xkblas_dposv_ir(int uplo, int N, int NRHS, const double* A, int LDA, const double* B, int LDB, double* X,
                     int LDX, int* iter);
/* CRT moduli of the reduced factor (default 7: about 24 bits per side at
 * inner dimension 4096; clamped to 4..13) and the correction limit (default
 * 30, as dsposv); values <= 0 (< 0 for itermax) restore the defaults. */
void xkblas_set_posv_ir(int moduli, int itermax);

/* device memory helpers for batched operands */
void* xkblas_device_malloc(size_t bytes);
void  xkblas_device_free(void* ptr);
int   xkblas_memcpy_htod(void* dst, const void* src, size_t bytes);
int   xkblas_memcpy_dtoh(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif
