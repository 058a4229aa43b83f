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

#ifdef __cplusplus
}
#endif

#endif
