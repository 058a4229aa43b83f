/*
 * xkblas_f77.h — Fortran-ABI entry points of libxkblas_amd.so.
 *
 * xkblas_d*_native_ forward to the host BLAS named by XKBLAS_BLASLIB
 * (reference blas/zxkblas_f77.h:38-107, blas/zgemm.c:383-396).  Integers
 * are LP64 `int`, as the reference's KBLAS_INT without MKL ILP64
 * (blas/common.gen:83-87).
 */
#ifndef XKBLAS_AMD_F77_H
#define XKBLAS_AMD_F77_H

#include "xkblas.h" /* Complex32_t, Complex64_t */

#ifdef __cplusplus
extern "C" {
#endif

void xkblas_dgemm_native_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
                          const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
                          const double* beta, double* C, const int* ldc);
void xkblas_dsyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                          const double* A, const int* lda, const double* beta, double* C, const int* ldc);
void xkblas_dtrsm_native_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m,
                          const int* n, const double* alpha, const double* A, const int* lda, double* B,
                          const int* ldb);
void xkblas_dpotrf_native_(const char* uplo, const int* n, double* A, const int* lda, int* info);
/* reference blas/zgemmt.c:402, zsymm.c:409, ztrmm.c:473, zsyr2k.c:339 */
void xkblas_dgemmt_native_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
                           const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
                           const double* beta, double* C, const int* ldc);
void xkblas_dsyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                           const double* A, const int* lda, const double* B, const int* ldb, const double* beta,
                           double* C, const int* ldc);
void xkblas_dsymm_native_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha,
                          const double* A, const int* lda, const double* B, const int* ldb, const double* beta,
                          double* C, const int* ldc);
void xkblas_dtrmm_native_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m,
                          const int* n, const double* alpha, const double* A, const int* lda, double* B,
                          const int* ldb);

/* complex forwarders (gen z, c of reference blas/zxkblas_f77.h) */
void xkblas_zgemm_native_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                          const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B,
                          const int* ldb, const Complex64_t* beta, Complex64_t* C, const int* ldc);
void xkblas_zgemmt_native_(const char* uplo, const char* ta, const char* tb, const int* n, const int* k,
                           const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B,
                           const int* ldb, const Complex64_t* beta, Complex64_t* C, const int* ldc);
void xkblas_ztrsm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B,
                          const int* ldb);
void xkblas_ztrmm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B,
                          const int* ldb);
void xkblas_zsymm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                          const Complex64_t* beta, Complex64_t* C, const int* ldc);
void xkblas_zhemm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                          const Complex64_t* beta, Complex64_t* C, const int* ldc);
void xkblas_zsyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* beta, Complex64_t* C,
                          const int* ldc);
void xkblas_zherk_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                          const Complex64_t* A, const int* lda, const double* beta, Complex64_t* C, const int* ldc);
void xkblas_zsyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                           const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                           const Complex64_t* beta, Complex64_t* C, const int* ldc);
void xkblas_zher2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                           const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                           const double* beta, Complex64_t* C, const int* ldc);
void xkblas_zpotrf_native_(const char* uplo, const int* n, Complex64_t* A, const int* lda, int* info);
void xkblas_cgemm_native_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                          const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B,
                          const int* ldb, const Complex32_t* beta, Complex32_t* C, const int* ldc);
void xkblas_cgemmt_native_(const char* uplo, const char* ta, const char* tb, const int* n, const int* k,
                           const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B,
                           const int* ldb, const Complex32_t* beta, Complex32_t* C, const int* ldc);
void xkblas_ctrsm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B,
                          const int* ldb);
void xkblas_ctrmm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B,
                          const int* ldb);
void xkblas_csymm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                          const Complex32_t* beta, Complex32_t* C, const int* ldc);
void xkblas_chemm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                          const Complex32_t* beta, Complex32_t* C, const int* ldc);
void xkblas_csyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* beta, Complex32_t* C,
                          const int* ldc);
void xkblas_cherk_native_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha,
                          const Complex32_t* A, const int* lda, const float* beta, Complex32_t* C, const int* ldc);
void xkblas_csyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                           const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                           const Complex32_t* beta, Complex32_t* C, const int* ldc);
void xkblas_cher2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                           const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                           const float* beta, Complex32_t* C, const int* ldc);
void xkblas_cpotrf_native_(const char* uplo, const int* n, Complex32_t* A, const int* lda, int* info);

#ifdef __cplusplus
}
#endif

#endif
