/*
 * xkblas.h — C API of the MI355X-native tiled BLAS-3 library (libxkblas_amd.so).
 *
 * Drop-in for the XKBlas C API on the ?gemm / ?syrk / ?trsm / ?potrf path.
 * Every declaration cites the reference interface it replaces (file:line in
 * the XKBlas sources).  Names, argument order, argument meaning and the CBLAS
 * enum values are those of the reference; only the implementation differs:
 * one GPU, an HBM tile cache, hand-written gfx950 kernels.
 *
 * Semantics shared by all *_async calls (reference blas/zgemm.c:118-370,
 * testing/ztesting_gemm.c:157-162):
 *   - matrices are column-major host arrays owned by the caller;
 *   - the call enqueues the work and returns; host copies of the outputs are
 *     valid only after xkblas_memory_coherent_async(...) + xkblas_sync();
 *   - device replicas are keyed by the host pointer and stay cached until
 *     xkblas_memory_invalidate[_caches]() (reference blas/xkblas.c:416-455);
 *   - an illegal argument prints a message and aborts (reference
 *     blas/common.gen:312-316).  With XKBLAS_ERROR_RETURN=1 in the environment
 *     the negative code listed by the reference is returned instead.
 */
#ifndef XKBLAS_AMD_XKBLAS_H
#define XKBLAS_AMD_XKBLAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CBLAS enum values, fixed by reference blas/xkblas.h:114-119. */
#ifndef KAAPI_NO_DEFAULT_BLAS_ENUM
typedef enum { CblasRowMajor = 101, CblasColMajor = 102 } CBLAS_ORDER;
typedef enum { CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113, CblasConjNoTrans = 114 } CBLAS_TRANSPOSE;
typedef enum { CblasUpper = 121, CblasLower = 122 } CBLAS_UPLO;
typedef enum { CblasNonUnit = 131, CblasUnit = 132 } CBLAS_DIAG;
typedef enum { CblasLeft = 141, CblasRight = 142 } CBLAS_SIDE;
typedef CBLAS_ORDER CBLAS_LAYOUT;
#endif

/* Complex element types (reference blas/xkblas.h:62-64): interleaved
 * (re, im), layout-compatible between C99 complex and std::complex. */
#ifdef __cplusplus
}
#include <complex>
typedef std::complex<float> Complex32_t;
typedef std::complex<double> Complex64_t;
extern "C" {
#else
#include <complex.h>
typedef float complex Complex32_t;
typedef double complex Complex64_t;
#endif
typedef double CFloat64_t;

/* Kernel list (reference blas/xkblas.h:124-139). */
typedef enum {
  KERN_VOID, KERN_GEMM, KERN_GEMMT, KERN_TRMM, KERN_TRSM, KERN_SYMM, KERN_SYRK,
  KERN_SYR2K, KERN_HEMM, KERN_HERK, KERN_HER2K, KERN_SWAP, KERN_COPYSCALE,
  KERN_POTRF /* extension: the reference has no POTRF kernel id */
} xkblas_kernel_t;

/* Math mode (reference blas/xkblas.h:100-103).  The two split modes are
 * extensions: FP64-accurate products computed from BF16 (resp. FP8) residue
 * matrices with exact FP32 accumulation and error-free CRT recombination.
 * XKBLAS_TENSOR_OP_MATH is accepted and behaves as the default (the reference
 * compiles it out on AMD, blas/ztask_gemm.c:182-196). */
typedef enum {
  XKBLAS_DEFAULT_MATH = 0,
  XKBLAS_TENSOR_OP_MATH = 1,
  XKBLAS_SPLIT_BF16_MATH = 2,
  XKBLAS_SPLIT_FP8_MATH = 3
} xkblas_mode_math_t;

/* ---- lifecycle and control ------------------------------------------------ */
int    xkblas_init(void);                      /* ref blas/xkblas.h:207, xkblas.c:1460 (ref-counted) */
int    xkblas_finalize(void);                  /* ref blas/xkblas.h:211, xkblas.c:1616 */
int    xkblas_sync(void);                      /* ref blas/xkblas.h:282, xkblas.c:1773 */
void   xkblas_set_param(size_t nb, size_t p);  /* ref blas/xkblas.h:484, xkblas.c:572 */
size_t xkblas_get_param(void);                 /* ref blas/xkblas.h:470 */
int    xkblas_get_ngpus(void);                 /* ref blas/xkblas.h:476 */
int    xkblas_set_ngpus(int ngpus);            /* ref blas/xkblas.h:480 (only 1 is accepted) */
int    xkblas_get_devicecount(void);           /* ref blas/xkblas.h:203 */
int    xkblas_get_device_count(int* count);    /* ref blas/xkblas.h:217 */
void   xkblas_set_modemath(xkblas_mode_math_t mode); /* ref blas/xkblas.h:107 */
const char* xkblas_get_configuration(void);    /* ref blas/xkblas.h:273 */
void   xkblas_load_sym(void** ptr, const char* name); /* ref blas/xkblas.h:278, xkblas.c:2678-2695 */
uint64_t xkblas_elapsedns(void);               /* ref blas/xkblas.h:492 */
double xkblas_elapsedtime(void);               /* ref blas/xkblas.h:498 */
/* ref blas/xkblas.h:310, xkblas.c:2239 */
size_t xkblas_auto_tilesize(void* xkctxt, xkblas_kernel_t kernel, size_t M, size_t N, size_t K);

/* Device properties (ref blas/xkblas.h:221-256, xkblas.c:1443).  The
 * reference's architecture codes are kept; MI355X, unknown to the reference
 * (its HIP plugin names MI250 and MI300A only, kaapi/kaapi_plugin_hip.c:
 * 1931-1937), gets the extension code XKBLAS_ARCH_MI355X.  has_unified is 0:
 * no unified memory is used.  Returns 0, or EINVAL (22) for a null pointer. */
typedef struct xkblas_dev_prop_t {
  int arch;
  int has_unified;
  int device_count;
} xkblas_dev_prop_t;
typedef enum xkblas_dev_arch_t {
  XKBLAS_ARCH_UNDEF = -1,
  XKBLAS_ARCH_UNKNOWN = 0,
  XKBLAS_ARCH_V100 = 1,
  XKBLAS_ARCH_A100 = 2,
  XKBLAS_ARCH_RESERVED0 = 3,
  XKBLAS_ARCH_GRACEHOPPER = 4,
  XKBLAS_ARCH_MI50 = 101,
  XKBLAS_ARCH_RESERVED1 = 102,
  XKBLAS_ARCH_RESERVED2 = 103,
  XKBLAS_ARCH_MI250 = 104,
  XKBLAS_ARCH_RESERVED3 = 105,
  XKBLAS_ARCH_MI300A = 106,
  XKBLAS_ARCH_MI355X = 107 /* extension */
} xkblas_dev_arch_t;
int    xkblas_get_device_properties(xkblas_dev_prop_t* prop); /* ref blas/xkblas.h:256 */
/* ref blas/xkblas.h:267: unified memory is not used, so EINVAL (22), the
 * reference's answer for devices without it (kaapi/kaapi_offload.c:631-640) */
int    xkblas_force_set_unified(void);

/* ---- memory and coherence ------------------------------------------------- */
void*  xkblas_malloc(size_t sz);               /* ref blas/xkblas.h:318, xkblas.c:622 (pinned) */
void   xkblas_free(void* ptr, size_t sz);      /* ref blas/xkblas.h:322 */
int    xkblas_memory_coherent_async(int uplo, int memflag, size_t M, size_t N,
                                    void* A, size_t ld, size_t elsize); /* ref blas/xkblas.h:441, xkblas.c:1953 */
int    xkblas_memory_syncall(void);            /* ref blas/xkblas.h:452 */
int    xkblas_memory_invalidate(const void* A);/* ref blas/xkblas.h:458 */
int    xkblas_memory_invalidate_caches(void);  /* ref blas/xkblas.h:462, xkblas.c:1855 */
int    xkblas_memory_free(void);               /* ref blas/xkblas.h:466 */
/* Host registration (ref blas/xkblas.h:323-366, xkblas.c:668-774).  Copies
 * from pageable memory already go through a ring of pinned staging buffers,
 * so registering existing memory has nothing to do: every request completes
 * at once with status 0 (test returns 1), handles are distinct and non-zero. */
uint64_t xkblas_register_memory_async(void* ptr, size_t sz);   /* ref blas/xkblas.h:328 */
uint64_t xkblas_unregister_memory_async(void* ptr, size_t sz); /* ref blas/xkblas.h:336 */
int    xkblas_register_memory_test(uint64_t handle);           /* ref blas/xkblas.h:345 */
int    xkblas_register_memory_wait(uint64_t handle);           /* ref blas/xkblas.h:351 */
int    xkblas_register_memory_waitall(void);                   /* ref blas/xkblas.h:355 */
int    xkblas_register_memory(void* ptr, size_t sz);           /* ref blas/xkblas.h:361 */
int    xkblas_unregister_memory(void* ptr, size_t sz);         /* ref blas/xkblas.h:365 */

/* ---- FP64 compute (gen d of reference blas/zxkblas.h) --------------------- */
int xkblas_dgemm_async(int transA, int transB, int M, int N, int K,
                       const double* alpha, const double* A, int LDA,
                       const double* B, int LDB,
                       const double* beta, double* C, int LDC);       /* ref blas/zxkblas.h:45, zgemm.c:118 */
int xkblas_dsyrk_async(int uplo, int trans, int N, int K,
                       const double* alpha, const double* A, int LDA,
                       const double* beta, double* C, int LDC);       /* ref blas/zxkblas.h:97, zsyrk.c:97 */
int xkblas_dtrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS,
                       const double* alpha, const double* A, int LDA,
                       double* B, int LDB);                           /* ref blas/zxkblas.h:69, ztrsm.c:102 */
/* Tile Cholesky.  The reference prototype (blas/zpotrf.c:46-48, unbuilt)
 * takes `const` A although it factors A in place; here A is non-const.
 * Returns 0 when enqueued; the LAPACK-style info is read after xkblas_sync()
 * with xkblas_potrf_info(A). */
int xkblas_dpotrf_async(int uplo, int N, double* A, int LDA);

/* Complex GEMM (gen z / c of reference blas/zxkblas.h:45, zgemm.c:118): same
 * arguments and codes; computed as one real GEMM of order (2M, N, 2K) on the
 * MFMA kernel (the split modes apply to the FP64 one). */
int xkblas_zgemm_async(int transA, int transB, int M, int N, int K,
                       const Complex64_t* alpha, const Complex64_t* A, int LDA,
                       const Complex64_t* B, int LDB,
                       const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_cgemm_async(int transA, int transB, int M, int N, int K,
                       const Complex32_t* alpha, const Complex32_t* A, int LDA,
                       const Complex32_t* B, int LDB,
                       const Complex32_t* beta, Complex32_t* C, int LDC);

/* Complex HERK and HEMM (reference blas/zxkblas.h:119,131, zherk.c, zhemm.c):
 * same arguments and codes (HERK's alpha, beta are real; the diagonal of C is
 * left real). */
int xkblas_zherk_async(int uplo, int trans, int N, int K,
                       const CFloat64_t* alpha, const Complex64_t* A, int LDA,
                       const CFloat64_t* beta, Complex64_t* C, int LDC);
int xkblas_cherk_async(int uplo, int trans, int N, int K,
                       const float* alpha, const Complex32_t* A, int LDA,
                       const float* beta, Complex32_t* C, int LDC);
int xkblas_zhemm_async(int side, int uplo, int M, int N,
                       const Complex64_t* alpha, const Complex64_t* A, int LDA,
                       const Complex64_t* B, int LDB,
                       const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_chemm_async(int side, int uplo, int M, int N,
                       const Complex32_t* alpha, const Complex32_t* A, int LDA,
                       const Complex32_t* B, int LDB,
                       const Complex32_t* beta, Complex32_t* C, int LDC);

/* The other complex L3 routines (reference blas/zxkblas.h gen z / c:
 * zsyrk.c, zsyr2k.c, zher2k.c, zgemmt.c, zsymm.c, ztrmm.c), same arguments
 * and codes; HER2K's beta is real. */
/* Complex Cholesky (A = L L^H or U^H U; reference blas/zpotrf.c:46-48, unbuilt
 * there); the LAPACK info is read after xkblas_sync() with xkblas_potrf_info. */
int xkblas_zpotrf_async(int uplo, int N, Complex64_t* A, int LDA);
int xkblas_cpotrf_async(int uplo, int N, Complex32_t* A, int LDA);
int xkblas_ztrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                       const Complex64_t* A, int LDA, Complex64_t* B, int LDB);   /* ref blas/zxkblas.h:69 */
int xkblas_ctrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                       const Complex32_t* A, int LDA, Complex32_t* B, int LDB);
int xkblas_zsyrk_async(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                       const Complex64_t* beta, Complex64_t* C, int LDC);   /* ref blas/zxkblas.h:97 */
/* ref blas/zxkblas.h:107 */
int xkblas_zsyr2k_async(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
/* ref blas/zxkblas.h:141 */
int xkblas_zher2k_async(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const double* beta, Complex64_t* C, int LDC);
/* ref blas/zxkblas.h:57 */
int xkblas_zgemmt_async(int uplo, int transA, int transB, int N, int K, const Complex64_t* alpha,
                        const Complex64_t* A, int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta,
                        Complex64_t* C, int LDC);
/* ref blas/zxkblas.h:85 */
int xkblas_zsymm_async(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                       const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_ztrmm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                       const Complex64_t* A, int LDA, Complex64_t* B, int LDB);   /* ref blas/zxkblas.h:77 */
int xkblas_csyrk_async(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                       const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_csyr2k_async(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_cher2k_async(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const float* beta, Complex32_t* C, int LDC);
int xkblas_cgemmt_async(int uplo, int transA, int transB, int N, int K, const Complex32_t* alpha,
                        const Complex32_t* A, int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta,
                        Complex32_t* C, int LDC);
int xkblas_csymm_async(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                       const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_ctrmm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                       const Complex32_t* A, int LDA, Complex32_t* B, int LDB);

/* The other L3 tile algorithms (reference blas/zgemmt.c, zsyr2k.c, zsymm.c,
 * ztrmm.c), same argument order and codes.  GEMMT writes only the uplo
 * triangle of C (all of C for uplo 0), as the host BLAS does. */
int xkblas_dgemmt_async(int uplo, int transA, int transB, int N, int K,
                        const double* alpha, const double* A, int LDA,
                        const double* B, int LDB,
                        const double* beta, double* C, int LDC);      /* ref blas/zxkblas.h:57, zgemmt.c:114 */
int xkblas_dsyr2k_async(int uplo, int trans, int N, int K,
                        const double* alpha, const double* A, int LDA,
                        const double* B, int LDB,
                        const double* beta, double* C, int LDC);      /* ref blas/zxkblas.h:107, zsyr2k.c:107 */
int xkblas_dsymm_async(int side, int uplo, int M, int N,
                       const double* alpha, const double* A, int LDA,
                       const double* B, int LDB,
                       const double* beta, double* C, int LDC);       /* ref blas/zxkblas.h:85, zsymm.c:107 */
int xkblas_dtrmm_async(int side, int uplo, int transA, int diag, int N, int NRHS,
                       const double* alpha, const double* A, int LDA,
                       double* B, int LDB);                           /* ref blas/zxkblas.h:77, ztrmm.c:101 */

/* Host-BLAS forwarders (reference blas/zgemm.c:373-414, zsyrk.c:322-357,
 * ztrsm.c:502-538): call the vendor BLAS named by XKBLAS_BLASLIB. */
int xkblas_dgemm_native(int transA, int transB, int M, int N, int K,
                        const double* alpha, const double* A, int LDA,
                        const double* B, int LDB,
                        const double* beta, double* C, int LDC);
int xkblas_dsyrk_native(int uplo, int trans, int N, int K,
                        const double* alpha, const double* A, int LDA,
                        const double* beta, double* C, int LDC);
int xkblas_dtrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS,
                        const double* alpha, const double* A, int LDA,
                        double* B, int LDB);
int xkblas_dpotrf_native(int uplo, int N, double* A, int LDA);
/* complex host-BLAS forwarders (gen z, c of the reference's *_native) */
int xkblas_zgemm_native(int transA, int transB, int M, int N, int K, const Complex64_t* alpha, const Complex64_t* A,
                        int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_zgemmt_native(int uplo, int transA, int transB, int N, int K, const Complex64_t* alpha,
                         const Complex64_t* A, int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta,
                         Complex64_t* C, int LDC);
int xkblas_ztrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                        const Complex64_t* A, int LDA, Complex64_t* B, int LDB);
int xkblas_ztrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                        const Complex64_t* A, int LDA, Complex64_t* B, int LDB);
int xkblas_zsymm_native(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_zhemm_native(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_zsyrk_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_zherk_native(int uplo, int trans, int N, int K, const double* alpha, const Complex64_t* A, int LDA,
                        const double* beta, Complex64_t* C, int LDC);
int xkblas_zsyr2k_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                         const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC);
int xkblas_zher2k_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                         const Complex64_t* B, int LDB, const double* beta, Complex64_t* C, int LDC);
int xkblas_zpotrf_native(int uplo, int N, Complex64_t* A, int LDA);
int xkblas_cgemm_native(int transA, int transB, int M, int N, int K, const Complex32_t* alpha, const Complex32_t* A,
                        int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_cgemmt_native(int uplo, int transA, int transB, int N, int K, const Complex32_t* alpha,
                         const Complex32_t* A, int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta,
                         Complex32_t* C, int LDC);
int xkblas_ctrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                        const Complex32_t* A, int LDA, Complex32_t* B, int LDB);
int xkblas_ctrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                        const Complex32_t* A, int LDA, Complex32_t* B, int LDB);
int xkblas_csymm_native(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_chemm_native(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_csyrk_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_cherk_native(int uplo, int trans, int N, int K, const float* alpha, const Complex32_t* A, int LDA,
                        const float* beta, Complex32_t* C, int LDC);
int xkblas_csyr2k_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                         const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC);
int xkblas_cher2k_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                         const Complex32_t* B, int LDB, const float* beta, Complex32_t* C, int LDC);
int xkblas_cpotrf_native(int uplo, int N, Complex32_t* A, int LDA);
int xkblas_dgemmt_native(int uplo, int transA, int transB, int N, int K,
                         const double* alpha, const double* A, int LDA,
                         const double* B, int LDB,
                         const double* beta, double* C, int LDC);   /* ref blas/zgemmt.c:432 */
int xkblas_dsyr2k_native(int uplo, int trans, int N, int K,
                         const double* alpha, const double* A, int LDA,
                         const double* B, int LDB,
                         const double* beta, double* C, int LDC);   /* ref blas/zxkblas.h:113 */
int xkblas_dsymm_native(int side, int uplo, int M, int N,
                        const double* alpha, const double* A, int LDA,
                        const double* B, int LDB,
                        const double* beta, double* C, int LDC);    /* ref blas/zsymm.c:426 */
int xkblas_dtrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS,
                        const double* alpha, const double* A, int LDA,
                        double* B, int LDB);                        /* ref blas/zxkblas.h:81 */

/* ---- FP32 compute (gen s of reference blas/zxkblas.h) --------------------- */
int xkblas_sgemm_async(int transA, int transB, int M, int N, int K,
                       const float* alpha, const float* A, int LDA,
                       const float* B, int LDB,
                       const float* beta, float* C, int LDC);
int xkblas_ssyrk_async(int uplo, int trans, int N, int K,
                       const float* alpha, const float* A, int LDA,
                       const float* beta, float* C, int LDC);
int xkblas_strsm_async(int side, int uplo, int transA, int diag, int N, int NRHS,
                       const float* alpha, const float* A, int LDA,
                       float* B, int LDB);
int xkblas_spotrf_async(int uplo, int N, float* A, int LDA);
int xkblas_sgemmt_async(int uplo, int transA, int transB, int N, int K,
                        const float* alpha, const float* A, int LDA,
                        const float* B, int LDB,
                        const float* beta, float* C, int LDC);
int xkblas_ssyr2k_async(int uplo, int trans, int N, int K,
                        const float* alpha, const float* A, int LDA,
                        const float* B, int LDB,
                        const float* beta, float* C, int LDC);
int xkblas_ssymm_async(int side, int uplo, int M, int N,
                       const float* alpha, const float* A, int LDA,
                       const float* B, int LDB,
                       const float* beta, float* C, int LDC);
int xkblas_strmm_async(int side, int uplo, int transA, int diag, int N, int NRHS,
                       const float* alpha, const float* A, int LDA,
                       float* B, int LDB);

/* ---- extensions (no reference counterpart) -------------------------------- */
/* LAPACK-style info of the last ?potrf_async on A (0 = success, k>0 = the
 * leading minor of order k is not positive definite, -1 = unknown A).  Valid
 * after xkblas_sync(). */
int xkblas_potrf_info(const void* A);

/* Number of CRT moduli used by the split modes (0 = choose from K and the
 * FP64 accuracy target).  Also read from XKBLAS_SPLIT_COUNT. */
void xkblas_set_split_count(int nsplit);
int  xkblas_get_split_count(void);

#ifdef __cplusplus
}
#endif

#include "xkblas_amd_ext.h"

#endif /* XKBLAS_AMD_XKBLAS_H */
