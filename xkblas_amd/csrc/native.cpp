// native.cpp — host-BLAS forwarders (the reference's *_native path).
//
// The reference resolves the vendor BLAS at run time: dlopen(XKBLAS_BLASLIB)
// then dlsym of the Fortran symbol (blas/xkblas.c:2678-2695,
// blas/zgemm.c:383-414).  The same is done here.  These calls run on the
// host cores and are the parity anchor of the reference's tests
// (testing/ztesting_gemm.c:262); they never touch the GPU.
#include "../../include/xkblas.h"
#include "../../include/xkblas_f77.h"
#include <dlfcn.h>
#include <mutex>
#include <cstdio>
#include <cstdlib>

namespace {

std::mutex g_sym_mu;
void* g_lib = nullptr;

const char* default_blaslib()
{
  const char* v = getenv("XKBLAS_BLASLIB");
  if (v && *v) return v;
  return "libmkl_rt.so";
}

void* open_lib()
{
  if (g_lib) return g_lib;
  const char* name = default_blaslib();
  g_lib = dlopen(name, RTLD_NOW | RTLD_LOCAL);
  if (!g_lib) {
    // MKL of the image lives outside the default search path
    char alt[512];
    snprintf(alt, sizeof alt, "/opt/conda/lib/%s", name);
    g_lib = dlopen(alt, RTLD_NOW | RTLD_LOCAL);
  }
  if (!g_lib) {
    fprintf(stderr, "[xkblas_amd] cannot load host BLAS '%s' (set XKBLAS_BLASLIB): %s\n", name, dlerror());
    abort();
  }
  return g_lib;
}

char trans_char(int t) { return t == CblasNoTrans ? 'N' : (t == CblasTrans ? 'T' : 'C'); }
char uplo_char(int u) { return u == CblasLower ? 'L' : 'U'; }

typedef void (*dgemm_t)(const char*, const char*, const int*, const int*, const int*, const double*, const double*,
                        const int*, const double*, const int*, const double*, double*, const int*);
typedef void (*dsyrk_t)(const char*, const char*, const int*, const int*, const double*, const double*, const int*,
                        const double*, double*, const int*);
typedef void (*dtrsm_t)(const char*, const char*, const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, double*, const int*);
typedef void (*dpotrf_t)(const char*, const int*, double*, const int*, int*);

typedef void (*dgemmt_t)(const char*, const char*, const char*, const int*, const int*, const double*, const double*,
                         const int*, const double*, const int*, const double*, double*, const int*);
typedef void (*dsyr2k_t)(const char*, const char*, const int*, const int*, const double*, const double*, const int*,
                         const double*, const int*, const double*, double*, const int*);
typedef void (*dsymm_t)(const char*, const char*, const int*, const int*, const double*, const double*, const int*,
                        const double*, const int*, const double*, double*, const int*);
typedef void (*dtrmm_t)(const char*, const char*, const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, double*, const int*);

dgemmt_t p_dgemmt = nullptr;
dsyr2k_t p_dsyr2k = nullptr;
dsymm_t p_dsymm = nullptr;
dtrmm_t p_dtrmm = nullptr;
dgemm_t p_dgemm = nullptr;
dsyrk_t p_dsyrk = nullptr;
dtrsm_t p_dtrsm = nullptr;
dpotrf_t p_dpotrf = nullptr;

}  // namespace

extern "C" {

void xkblas_load_sym(void** ptr, const char* name)
{
  std::lock_guard<std::mutex> l(g_sym_mu);
  void* lib = open_lib();
  void* s = dlsym(lib, name);
  if (!s) {
    fprintf(stderr, "[xkblas_amd] symbol '%s' not found in host BLAS\n", name);
    abort();
  }
  *ptr = s;
}

void xkblas_dgemm_native_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
                          const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
                          const double* beta, double* C, const int* ldc)
{
  if (!p_dgemm) xkblas_load_sym((void**)&p_dgemm, "dgemm_");
  p_dgemm(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void xkblas_dsyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                          const double* A, const int* lda, const double* beta, double* C, const int* ldc)
{
  if (!p_dsyrk) xkblas_load_sym((void**)&p_dsyrk, "dsyrk_");
  p_dsyrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

void xkblas_dtrsm_native_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m,
                          const int* n, const double* alpha, const double* A, const int* lda, double* B,
                          const int* ldb)
{
  if (!p_dtrsm) xkblas_load_sym((void**)&p_dtrsm, "dtrsm_");
  p_dtrsm(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}

void xkblas_dpotrf_native_(const char* uplo, const int* n, double* A, const int* lda, int* info)
{
  if (!p_dpotrf) xkblas_load_sym((void**)&p_dpotrf, "dpotrf_");
  p_dpotrf(uplo, n, A, lda, info);
}

// reference blas/zgemmt.c:402-446, zsyr2k.c:339-369, zsymm.c:409-440, ztrmm.c:473-500
void xkblas_dgemmt_native_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
                           const double* alpha, const double* A, const int* lda, const double* B, const int* ldb,
                           const double* beta, double* C, const int* ldc)
{
  if (!p_dgemmt) xkblas_load_sym((void**)&p_dgemmt, "dgemmt_");
  p_dgemmt(uplo, transa, transb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void xkblas_dsyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                           const double* A, const int* lda, const double* B, const int* ldb, const double* beta,
                           double* C, const int* ldc)
{
  if (!p_dsyr2k) xkblas_load_sym((void**)&p_dsyr2k, "dsyr2k_");
  p_dsyr2k(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void xkblas_dsymm_native_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha,
                          const double* A, const int* lda, const double* B, const int* ldb, const double* beta,
                          double* C, const int* ldc)
{
  if (!p_dsymm) xkblas_load_sym((void**)&p_dsymm, "dsymm_");
  p_dsymm(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

void xkblas_dtrmm_native_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m,
                          const int* n, const double* alpha, const double* A, const int* lda, double* B,
                          const int* ldb)
{
  if (!p_dtrmm) xkblas_load_sym((void**)&p_dtrmm, "dtrmm_");
  p_dtrmm(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}

int xkblas_dgemmt_native(int uplo, int transA, int transB, int N, int K, const double* alpha, const double* A,
                         int LDA, const double* B, int LDB, const double* beta, double* C, int LDC)
{
  char u = uplo_char(uplo), ta = trans_char(transA), tb = trans_char(transB);
  xkblas_dgemmt_native_(&u, &ta, &tb, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

int xkblas_dsyr2k_native(int uplo, int trans, int N, int K, const double* alpha, const double* A, int LDA,
                         const double* B, int LDB, const double* beta, double* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_dsyr2k_native_(&u, &t, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

int xkblas_dsymm_native(int side, int uplo, int M, int N, const double* alpha, const double* A, int LDA,
                        const double* B, int LDB, const double* beta, double* C, int LDC)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo);
  xkblas_dsymm_native_(&s, &u, &M, &N, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

int xkblas_dtrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const double* alpha,
                        const double* A, int LDA, double* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), t = trans_char(transA);
  char d = diag == CblasUnit ? 'U' : 'N';
  xkblas_dtrmm_native_(&s, &u, &t, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

int xkblas_dgemm_native(int transA, int transB, int M, int N, int K, const double* alpha, const double* A, int LDA,
                        const double* B, int LDB, const double* beta, double* C, int LDC)
{
  char ta = trans_char(transA), tb = trans_char(transB);
  xkblas_dgemm_native_(&ta, &tb, &M, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

int xkblas_dsyrk_native(int uplo, int trans, int N, int K, const double* alpha, const double* A, int LDA,
                        const double* beta, double* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_dsyrk_native_(&u, &t, &N, &K, alpha, A, &LDA, beta, C, &LDC);
  return 0;
}

int xkblas_dtrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const double* alpha,
                        const double* A, int LDA, double* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), t = trans_char(transA);
  char d = diag == CblasUnit ? 'U' : 'N';
  xkblas_dtrsm_native_(&s, &u, &t, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

int xkblas_dpotrf_native(int uplo, int N, double* A, int LDA)
{
  char u = uplo_char(uplo);
  int info = 0;
  xkblas_dpotrf_native_(&u, &N, A, &LDA, &info);
  return info;
}

}  // extern "C"
