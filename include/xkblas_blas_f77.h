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

#include "xkblas.h" /* Complex32_t, Complex64_t */

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

/* the other precisions (reference blas/zlibxkblas_wrapper.c:46-400, gen s, c, z) */
void zgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
            const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
            const Complex64_t* beta, Complex64_t* C, const int* ldc);
void zgemmt_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
             const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
             const Complex64_t* beta, Complex64_t* C, const int* ldc);
void ztrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B, const int* ldb);
void ztrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B, const int* ldb);
void zsymm_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
            const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb, const Complex64_t* beta,
            Complex64_t* C, const int* ldc);
void zsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
            const Complex64_t* A, const int* lda, const Complex64_t* beta, Complex64_t* C, const int* ldc);
void zsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
             const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb, const Complex64_t* beta,
             Complex64_t* C, const int* ldc);
void zhemm_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
            const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb, const Complex64_t* beta,
            Complex64_t* C, const int* ldc);
void zherk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
            const Complex64_t* A, const int* lda, const double* beta, Complex64_t* C, const int* ldc);
void zher2k_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
             const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb, const double* beta,
             Complex64_t* C, const int* ldc);
void cgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
            const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
            const Complex32_t* beta, Complex32_t* C, const int* ldc);
void cgemmt_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
             const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
             const Complex32_t* beta, Complex32_t* C, const int* ldc);
void ctrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B, const int* ldb);
void ctrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B, const int* ldb);
void csymm_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
            const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb, const Complex32_t* beta,
            Complex32_t* C, const int* ldc);
void csyrk_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
            const Complex32_t* A, const int* lda, const Complex32_t* beta, Complex32_t* C, const int* ldc);
void csyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
             const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb, const Complex32_t* beta,
             Complex32_t* C, const int* ldc);
void chemm_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
            const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb, const Complex32_t* beta,
            Complex32_t* C, const int* ldc);
void cherk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha,
            const Complex32_t* A, const int* lda, const float* beta, Complex32_t* C, const int* ldc);
void cher2k_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
             const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb, const float* beta,
             Complex32_t* C, const int* ldc);
void ssyrk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const float* A,
            const int* lda, const float* beta, float* C, const int* ldc);
void strsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const float* alpha, const float* A, const int* lda, float* B, const int* ldb);

#ifdef __cplusplus
}
#endif

#endif
