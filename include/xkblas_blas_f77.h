/*
 * xkblas_blas_f77.h — the Fortran BLAS/LAPACK symbols exported by
 * libxkblas_amd_wrapper.so (LD_PRELOAD or link-time drop-in for a host
 * BLAS).  Reference: the unbuilt wrapper blas/libxkblas_wrapper.c:88-187 and
 * blas/zlibxkblas_wrapper.c:46-76, 122-148, 227-252; dpotrf_ follows LAPACK
 * (no reference counterpart).  Each call is synchronous; it runs on the host
 * BLAS when its arithmetic intensity is below XKBLAS_THRESHOLD[_<KERNEL>]
 * (default 1600, as in the reference), else on the GPU.
 */
#ifndef XKBLAS_AMD_BLAS_F77_H
#define XKBLAS_AMD_BLAS_F77_H

#ifdef __cplusplus
extern "C" {
#endif

void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const double* alpha,
            const double* A, const int* lda, const double* B, const int* ldb, const double* beta, double* C,
            const int* ldc);
void sgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
            const float* A, const int* lda, const float* B, const int* ldb, const float* beta, float* C,
            const int* ldc);
void dsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* A,
            const int* lda, const double* beta, double* C, const int* ldc);
void dtrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* A, const int* lda, double* B, const int* ldb);
void dpotrf_(const char* uplo, const int* n, double* A, const int* lda, int* info);
/* reference blas/zlibxkblas_wrapper.c:85-300 (gemmt, trmm, symm, syr2k) */
void dgemmt_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
             const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
             const double* beta, double* C, const int* ldc);
void dsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* A,
             const int* lda, const double* B, const int* ldb, const double* beta, double* C, const int* ldc);
void dsymm_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha, const double* A,
            const int* lda, const double* B, const int* ldb, const double* beta, double* C, const int* ldc);
void dtrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* A, const int* lda, double* B, const int* ldb);

#ifdef __cplusplus
}
#endif

#endif
