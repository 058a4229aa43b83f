// wrapper.cpp — libxkblas_amd_wrapper.so: Fortran-ABI drop-in dgemm_,
// dsyrk_, dtrsm_, dpotrf_, sgemm_, and dgemmt_, dsyr2k_, dsymm_, dtrmm_
// (blas/zlibxkblas_wrapper.c:85-300) for LD_PRELOAD or link-time
// substitution of a host BLAS.
//
// Restates the reference's unbuilt wrapper (blas/libxkblas_wrapper.c:88-187,
// blas/zlibxkblas_wrapper.c:46-76): quick return, then a CPU/GPU choice on
// the arithmetic intensity flops/bytes of the call (blas/flops.h FLOPS_* /
// DATA_*) against XKBLAS_THRESHOLD (default 1600, per kernel
// XKBLAS_THRESHOLD_DGEMM, ...).  Below it the host BLAS of XKBLAS_BLASLIB
// runs the call; above it the GPU path runs synchronously (async + coherent +
// sync + invalidate).  dpotrf_ has no counterpart in the reference; it
// follows LAPACK's interface.
#include "../../include/xkblas.h"
#include "../../include/xkblas_f77.h"
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

enum Kern { K_DGEMM, K_SGEMM, K_DSYRK, K_SSYRK, K_DTRSM, K_STRSM, K_DPOTRF, K_DGEMMT, K_DSYR2K, K_DSYMM, K_DTRMM,
            K_LAST };
const char* kKernName[K_LAST] = {"DGEMM", "SGEMM", "DSYRK", "SSYRK", "DTRSM", "STRSM", "DPOTRF",
                                 "DGEMMT", "DSYR2K", "DSYMM", "DTRMM"};
double g_threshold[K_LAST];
bool g_gpu_ready = false;

// GPU side initialised on the first call that goes to the GPU, so that a
// process whose calls all stay on the host never touches the device (the
// reference initialises in the library constructor).
void wrapper_exit() { xkblas_finalize(); }

void gpu_ready()
{
  if (g_gpu_ready) return;
  if (xkblas_init() != 0) {
    fprintf(stderr, "libxkblas_amd_wrapper: cannot initialize\n");
    abort();
  }
  // registered after the HIP runtime's own exit handlers, so it runs before
  // them (a library destructor would run after the runtime is gone)
  atexit(wrapper_exit);
  g_gpu_ready = true;
}

__attribute__((constructor)) void wrapper_init()
{
  long th = 1600;
  if (const char* v = getenv("XKBLAS_THRESHOLD")) th = atol(v) < 0 ? 0 : atol(v);
  for (int i = 0; i < K_LAST; ++i) {
    char name[64];
    snprintf(name, sizeof name, "XKBLAS_THRESHOLD_%s", kKernName[i]);
    const char* v = getenv(name);
    g_threshold[i] = v ? (atol(v) < 0 ? 0.0 : (double)atol(v)) : (double)th;
  }
}

int trans_of(const char* t)
{
  switch (t[0]) {
    case 'n': case 'N': return CblasNoTrans;
    case 't': case 'T': return CblasTrans;
    case 'c': case 'C': return CblasConjTrans;
    default: abort();
  }
}
int uplo_of(const char* u) { return (u[0] == 'l' || u[0] == 'L') ? CblasLower : CblasUpper; }
int side_of(const char* s) { return (s[0] == 'l' || s[0] == 'L') ? CblasLeft : CblasRight; }
int diag_of(const char* d) { return (d[0] == 'u' || d[0] == 'U') ? CblasUnit : CblasNonUnit; }

// arithmetic intensity of the reference's flop/byte model (blas/flops.h)
double ai_gemm(double m, double n, double k, double elt) { return 2.0 * m * n * k / (elt * (m * n + m * k + k * n)); }
double ai_syrk(double n, double k, double elt) { return k * n * (n + 1.0) / (elt * (0.5 * n * n + n * k)); }
double ai_trsm(bool left, double m, double n, double elt)
{
  const double f = left ? n * m * (m + 1.0) : m * n * (n + 1.0);
  const double d = left ? (0.5 * m * m + 2.0 * m * n) : (0.5 * n * n + 2.0 * n * m);
  return f / (elt * d);
}

// FLOPS_DGEMMT / DATA_DGEMMT, FLOPS_DSYR2K / DATA_DSYR2K, FLOPS_DSYMM /
// DATA_DSYMM, FLOPS_DTRMM / DATA_DTRMM (blas/flops.h:55-75,375-417)
double ai_gemmt(double n, double k, double elt) { return n * n * k / (elt * (0.5 * n * n + 2.0 * n * k)); }
double ai_syr2k(double n, double k, double elt) { return (2.0 * k * n * n + n) / (elt * (0.5 * n * n + 2.0 * n * k)); }
double ai_symm(bool left, double m, double n, double elt)
{
  return left ? 2.0 * m * m * n / (elt * (m * m + 2.0 * m * n)) : 2.0 * m * n * n / (elt * (n * n + 2.0 * m * n));
}
double ai_trmm(bool left, double m, double n, double elt)
{
  const double f = left ? n * m * m : m * n * n;
  const double d = left ? (0.5 * m * m + 2.0 * m * n) : (0.5 * n * n + 2.0 * n * m);
  return f / (elt * d);
}

// zero / one tests of a scalar argument, real or complex
template <typename T>
bool is0(const T* v)
{
  return *v == T(0);
}
template <typename T>
bool is1(const T* v)
{
  return *v == T(1);
}

template <typename T>
void coherent_sync(int uplo, int m, int n, T* C, int ldc)
{
  xkblas_memory_coherent_async(uplo, 0, m, n, C, ldc, sizeof(T));
  xkblas_sync();
  xkblas_memory_invalidate_caches();
}

}  // namespace

extern "C" {

void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const double* alpha,
            const double* A, const int* lda, const double* B, const int* ldb, const double* beta, double* C,
            const int* ldc)
{
  if (*m == 0 || *n == 0 || ((*alpha == 0.0 || *k == 0) && *beta == 1.0)) return;
  if (ai_gemm(*m, *n, *k, 8.0) < g_threshold[K_DGEMM]) {
    xkblas_dgemm_native_(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_dgemm_async(trans_of(transa), trans_of(transb), *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
  coherent_sync<double>(0, *m, *n, C, *ldc);
}

void sgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
            const float* A, const int* lda, const float* B, const int* ldb, const float* beta, float* C,
            const int* ldc)
{
  if (*m == 0 || *n == 0 || ((*alpha == 0.0f || *k == 0) && *beta == 1.0f)) return;
  if (ai_gemm(*m, *n, *k, 4.0) < g_threshold[K_SGEMM]) {
    typedef void (*f_t)(const char*, const char*, const int*, const int*, const int*, const float*, const float*,
                        const int*, const float*, const int*, const float*, float*, const int*);
    static f_t f = nullptr;
    if (!f) xkblas_load_sym((void**)&f, "sgemm_");
    f(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_sgemm_async(trans_of(transa), trans_of(transb), *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
  coherent_sync<float>(0, *m, *n, C, *ldc);
}

void dsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* A,
            const int* lda, const double* beta, double* C, const int* ldc)
{
  if (*n == 0 || ((*alpha == 0.0 || *k == 0) && *beta == 1.0)) return;
  if (ai_syrk(*n, *k, 8.0) < g_threshold[K_DSYRK]) {
    xkblas_dsyrk_native_(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_dsyrk_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, beta, C, *ldc);
  coherent_sync<double>(uplo_of(uplo), *n, *n, C, *ldc);
}

void dtrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* A, const int* lda, double* B, const int* ldb)
{
  if (*m == 0 || *n == 0) return;
  if (ai_trsm(side_of(side) == CblasLeft, *m, *n, 8.0) < g_threshold[K_DTRSM]) {
    xkblas_dtrsm_native_(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
    return;
  }
  gpu_ready();
  xkblas_dtrsm_async(side_of(side), uplo_of(uplo), trans_of(transa), diag_of(diag), *m, *n, alpha, A, *lda, B,
                     *ldb);
  coherent_sync<double>(0, *m, *n, B, *ldb);
}

void dgemmt_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
             const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
             const double* beta, double* C, const int* ldc)
{
  if (*n == 0 || ((*alpha == 0.0 || *k == 0) && *beta == 1.0)) return;
  if (ai_gemmt(*n, *k, 8.0) < g_threshold[K_DGEMMT]) {
    xkblas_dgemmt_native_(uplo, transa, transb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_dgemmt_async(uplo_of(uplo), trans_of(transa), trans_of(transb), *n, *k, alpha, A, *lda, B, *ldb, beta, C,
                      *ldc);
  coherent_sync<double>(uplo_of(uplo), *n, *n, C, *ldc);
}

void dsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* A,
             const int* lda, const double* B, const int* ldb, const double* beta, double* C, const int* ldc)
{
  if (*n == 0 || ((*alpha == 0.0 || *k == 0) && *beta == 1.0)) return;
  if (ai_syr2k(*n, *k, 8.0) < g_threshold[K_DSYR2K]) {
    xkblas_dsyr2k_native_(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_dsyr2k_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
  coherent_sync<double>(uplo_of(uplo), *n, *n, C, *ldc);
}

void dsymm_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha, const double* A,
            const int* lda, const double* B, const int* ldb, const double* beta, double* C, const int* ldc)
{
  if (*m == 0 || *n == 0 || (*alpha == 0.0 && *beta == 1.0)) return;
  if (ai_symm(side_of(side) == CblasLeft, *m, *n, 8.0) < g_threshold[K_DSYMM]) {
    xkblas_dsymm_native_(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  gpu_ready();
  xkblas_dsymm_async(side_of(side), uplo_of(uplo), *m, *n, alpha, A, *lda, B, *ldb, beta, C, *ldc);
  coherent_sync<double>(0, *m, *n, C, *ldc);
}

void dtrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* A, const int* lda, double* B, const int* ldb)
{
  if (*m == 0 || *n == 0) return;
  if (ai_trmm(side_of(side) == CblasLeft, *m, *n, 8.0) < g_threshold[K_DTRMM]) {
    xkblas_dtrmm_native_(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
    return;
  }
  gpu_ready();
  xkblas_dtrmm_async(side_of(side), uplo_of(uplo), trans_of(transa), diag_of(diag), *m, *n, alpha, A, *lda, B,
                     *ldb);
  coherent_sync<double>(0, *m, *n, B, *ldb);
}

void dpotrf_(const char* uplo, const int* n, double* A, const int* lda, int* info)
{
  *info = 0;
  if (*n == 0) return;
  const double ai = (*n / 3.0) / (8.0 * 0.5);  // (n^3/3) / (8 n^2 / 2)
  if (ai < g_threshold[K_DPOTRF]) {
    xkblas_dpotrf_native_(uplo, n, A, lda, info);
    return;
  }
  gpu_ready();
  if (xkblas_dpotrf_async(uplo_of(uplo), *n, A, *lda) != 0) {
    *info = -1;
    return;
  }
  xkblas_memory_coherent_async(uplo_of(uplo), 0, *n, *n, A, *lda, sizeof(double));
  xkblas_sync();
  *info = xkblas_potrf_info(A);
  xkblas_memory_invalidate_caches();
}

// ---- the other precisions of blas/zlibxkblas_wrapper.c:46-400 (s, c, z):
// same dispatch; the host branch calls the host BLAS routine of the same
// name.  Complex arithmetic intensity uses 4 real flops per complex
// multiply-add pair and 2x the element size.
#define XKW_HOST(F, NAME, ...)                      \
  do {                                              \
    static F f_ = nullptr;                          \
    if (!f_) xkblas_load_sym((void**)&f_, NAME);    \
    f_(__VA_ARGS__);                                \
  } while (0)

#define XKW_THRESHOLD(KEY)                                                            \
  static const double thr_ = [] {                                                     \
    const char* v = getenv("XKBLAS_THRESHOLD_" KEY);                                  \
    if (v) return atol(v) < 0 ? 0.0 : (double)atol(v);                                \
    const char* g = getenv("XKBLAS_THRESHOLD");                                       \
    return g ? (atol(g) < 0 ? 0.0 : (double)atol(g)) : 1600.0;                        \
  }()

#define XKW_FAMILY(P, KEYP, S, R, CPLX)                                                                          \
  void P##gemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const S* alpha,        \
                const S* A, const int* lda, const S* B, const int* ldb, const S* beta, S* C, const int* ldc)     \
  {                                                                                                             \
    if (*m == 0 || *n == 0 || ((is0(alpha) || *k == 0) && is1(beta))) return;                                    \
    XKW_THRESHOLD(KEYP "GEMM");                                                                                 \
    if ((CPLX ? 4.0 : 1.0) * ai_gemm(*m, *n, *k, sizeof(S)) < thr_) {                                           \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const int*, const S*, const S*,       \
                          const int*, const S*, const int*, const S*, S*, const int*);                          \
      XKW_HOST(f_t, #P "gemm_", ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);                          \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##gemm_async(trans_of(ta), trans_of(tb), *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);       \
    coherent_sync<S>(0, *m, *n, C, *ldc);                                                                       \
  }                                                                                                             \
  void P##gemmt_(const char* uplo, const char* ta, const char* tb, const int* n, const int* k, const S* alpha,   \
                 const S* A, const int* lda, const S* B, const int* ldb, const S* beta, S* C, const int* ldc)    \
  {                                                                                                             \
    if (*n == 0 || ((is0(alpha) || *k == 0) && is1(beta))) return;                                              \
    XKW_THRESHOLD(KEYP "GEMMT");                                                                                \
    if ((CPLX ? 4.0 : 1.0) * ai_gemmt(*n, *k, sizeof(S)) < thr_) {                                              \
      typedef void (*f_t)(const char*, const char*, const char*, const int*, const int*, const S*, const S*,      \
                          const int*, const S*, const int*, const S*, S*, const int*);                          \
      XKW_HOST(f_t, #P "gemmt_", uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);                      \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##gemmt_async(uplo_of(uplo), trans_of(ta), trans_of(tb), *n, *k, alpha, A, *lda, B, *ldb, beta, C, \
                            *ldc);                                                                              \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }                                                                                                             \
  void P##trsm_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m, const int* n, \
                const S* alpha, const S* A, const int* lda, S* B, const int* ldb)                                \
  {                                                                                                             \
    if (*m == 0 || *n == 0) return;                                                                             \
    XKW_THRESHOLD(KEYP "TRSM");                                                                                 \
    if ((CPLX ? 4.0 : 1.0) * ai_trsm(side_of(side) == CblasLeft, *m, *n, sizeof(S)) < thr_) {                   \
      typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const S*,   \
                          const S*, const int*, S*, const int*);                                                \
      XKW_HOST(f_t, #P "trsm_", side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);                             \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##trsm_async(side_of(side), uplo_of(uplo), trans_of(ta), diag_of(diag), *m, *n, alpha, A, *lda, B,  \
                           *ldb);                                                                               \
    coherent_sync<S>(0, *m, *n, B, *ldb);                                                                       \
  }                                                                                                             \
  void P##trmm_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m, const int* n, \
                const S* alpha, const S* A, const int* lda, S* B, const int* ldb)                                \
  {                                                                                                             \
    if (*m == 0 || *n == 0) return;                                                                             \
    XKW_THRESHOLD(KEYP "TRMM");                                                                                 \
    if ((CPLX ? 4.0 : 1.0) * ai_trmm(side_of(side) == CblasLeft, *m, *n, sizeof(S)) < thr_) {                   \
      typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const S*,   \
                          const S*, const int*, S*, const int*);                                                \
      XKW_HOST(f_t, #P "trmm_", side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);                             \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##trmm_async(side_of(side), uplo_of(uplo), trans_of(ta), diag_of(diag), *m, *n, alpha, A, *lda, B,  \
                           *ldb);                                                                               \
    coherent_sync<S>(0, *m, *n, B, *ldb);                                                                       \
  }                                                                                                             \
  void P##symm_(const char* side, const char* uplo, const int* m, const int* n, const S* alpha, const S* A,       \
                const int* lda, const S* B, const int* ldb, const S* beta, S* C, const int* ldc)                 \
  {                                                                                                             \
    if (*m == 0 || *n == 0 || (is0(alpha) && is1(beta))) return;                                                \
    XKW_THRESHOLD(KEYP "SYMM");                                                                                 \
    if ((CPLX ? 4.0 : 1.0) * ai_symm(side_of(side) == CblasLeft, *m, *n, sizeof(S)) < thr_) {                   \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, const int*, const S*, S*, const int*);                                      \
      XKW_HOST(f_t, #P "symm_", side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);                         \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##symm_async(side_of(side), uplo_of(uplo), *m, *n, alpha, A, *lda, B, *ldb, beta, C, *ldc);        \
    coherent_sync<S>(0, *m, *n, C, *ldc);                                                                       \
  }                                                                                                             \
  void P##syrk_(const char* uplo, const char* trans, const int* n, const int* k, const S* alpha, const S* A,     \
                const int* lda, const S* beta, S* C, const int* ldc)                                            \
  {                                                                                                             \
    if (*n == 0 || ((is0(alpha) || *k == 0) && is1(beta))) return;                                              \
    XKW_THRESHOLD(KEYP "SYRK");                                                                                 \
    if ((CPLX ? 4.0 : 1.0) * ai_syrk(*n, *k, sizeof(S)) < thr_) {                                               \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, S*, const int*);                                                            \
      XKW_HOST(f_t, #P "syrk_", uplo, trans, n, k, alpha, A, lda, beta, C, ldc);                                \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##syrk_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, beta, C, *ldc);              \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }                                                                                                             \
  void P##syr2k_(const char* uplo, const char* trans, const int* n, const int* k, const S* alpha, const S* A,    \
                 const int* lda, const S* B, const int* ldb, const S* beta, S* C, const int* ldc)                \
  {                                                                                                             \
    if (*n == 0 || ((is0(alpha) || *k == 0) && is1(beta))) return;                                              \
    XKW_THRESHOLD(KEYP "SYR2K");                                                                                \
    if ((CPLX ? 4.0 : 1.0) * ai_syr2k(*n, *k, sizeof(S)) < thr_) {                                              \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, const int*, const S*, S*, const int*);                                      \
      XKW_HOST(f_t, #P "syr2k_", uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);                       \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##syr2k_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);     \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }

#define XKW_HERM(P, KEYP, S, R)                                                                                  \
  void P##hemm_(const char* side, const char* uplo, const int* m, const int* n, const S* alpha, const S* A,       \
                const int* lda, const S* B, const int* ldb, const S* beta, S* C, const int* ldc)                 \
  {                                                                                                             \
    if (*m == 0 || *n == 0 || (is0(alpha) && is1(beta))) return;                                                \
    XKW_THRESHOLD(KEYP "HEMM");                                                                                 \
    if (4.0 * ai_symm(side_of(side) == CblasLeft, *m, *n, sizeof(S)) < thr_) {                                  \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, const int*, const S*, S*, const int*);                                      \
      XKW_HOST(f_t, #P "hemm_", side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);                         \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##hemm_async(side_of(side), uplo_of(uplo), *m, *n, alpha, A, *lda, B, *ldb, beta, C, *ldc);        \
    coherent_sync<S>(0, *m, *n, C, *ldc);                                                                       \
  }                                                                                                             \
  void P##herk_(const char* uplo, const char* trans, const int* n, const int* k, const R* alpha, const S* A,     \
                const int* lda, const R* beta, S* C, const int* ldc)                                            \
  {                                                                                                             \
    if (*n == 0 || ((*alpha == R(0) || *k == 0) && *beta == R(1))) return;                                       \
    XKW_THRESHOLD(KEYP "HERK");                                                                                 \
    if (4.0 * ai_syrk(*n, *k, sizeof(S)) < thr_) {                                                              \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const R*, const S*, const int*,       \
                          const R*, S*, const int*);                                                            \
      XKW_HOST(f_t, #P "herk_", uplo, trans, n, k, alpha, A, lda, beta, C, ldc);                                \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##herk_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, beta, C, *ldc);              \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }                                                                                                             \
  void P##her2k_(const char* uplo, const char* trans, const int* n, const int* k, const S* alpha, const S* A,    \
                 const int* lda, const S* B, const int* ldb, const R* beta, S* C, const int* ldc)                \
  {                                                                                                             \
    if (*n == 0 || ((is0(alpha) || *k == 0) && *beta == R(1))) return;                                          \
    XKW_THRESHOLD(KEYP "HER2K");                                                                                \
    if (4.0 * ai_syr2k(*n, *k, sizeof(S)) < thr_) {                                                             \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, const int*, const R*, S*, const int*);                                      \
      XKW_HOST(f_t, #P "her2k_", uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);                       \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##her2k_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);     \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }

XKW_FAMILY(z, "Z", Complex64_t, double, true)
XKW_FAMILY(c, "C", Complex32_t, float, true)
XKW_HERM(z, "Z", Complex64_t, double)
XKW_HERM(c, "C", Complex32_t, float)

// single precision: the entries dgemm_ ... have no s twin above yet
#define XKW_SREAL(P, KEYP, S)                                                                                    \
  void P##syrk_(const char* uplo, const char* trans, const int* n, const int* k, const S* alpha, const S* A,     \
                const int* lda, const S* beta, S* C, const int* ldc)                                            \
  {                                                                                                             \
    if (*n == 0 || ((is0(alpha) || *k == 0) && is1(beta))) return;                                              \
    XKW_THRESHOLD(KEYP "SYRK");                                                                                 \
    if (ai_syrk(*n, *k, sizeof(S)) < thr_) {                                                                    \
      typedef void (*f_t)(const char*, const char*, const int*, const int*, const S*, const S*, const int*,       \
                          const S*, S*, const int*);                                                            \
      XKW_HOST(f_t, #P "syrk_", uplo, trans, n, k, alpha, A, lda, beta, C, ldc);                                \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##syrk_async(uplo_of(uplo), trans_of(trans), *n, *k, alpha, A, *lda, beta, C, *ldc);              \
    coherent_sync<S>(uplo_of(uplo), *n, *n, C, *ldc);                                                           \
  }                                                                                                             \
  void P##trsm_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m, const int* n, \
                const S* alpha, const S* A, const int* lda, S* B, const int* ldb)                                \
  {                                                                                                             \
    if (*m == 0 || *n == 0) return;                                                                             \
    XKW_THRESHOLD(KEYP "TRSM");                                                                                 \
    if (ai_trsm(side_of(side) == CblasLeft, *m, *n, sizeof(S)) < thr_) {                                        \
      typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const S*,   \
                          const S*, const int*, S*, const int*);                                                \
      XKW_HOST(f_t, #P "trsm_", side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);                             \
      return;                                                                                                   \
    }                                                                                                           \
    gpu_ready();                                                                                                \
    xkblas_##P##trsm_async(side_of(side), uplo_of(uplo), trans_of(ta), diag_of(diag), *m, *n, alpha, A, *lda, B,  \
                           *ldb);                                                                               \
    coherent_sync<S>(0, *m, *n, B, *ldb);                                                                       \
  }
XKW_SREAL(s, "S", float)

}  // extern "C"
