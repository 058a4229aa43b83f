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

// ---- complex forwarders (reference blas/z*.c *_native / *_native_, gen z, c)
void xkblas_zgemm_native_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                          const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B,
                          const int* ldb, const Complex64_t* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const int*, const Complex64_t*,
                      const Complex64_t*, const int*, const Complex64_t*, const int*, const Complex64_t*, Complex64_t*,
                const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zgemm_");
  f(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zgemm_native(int transA, int transB, int M, int N, int K, const Complex64_t* alpha, const Complex64_t* A,
                        int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  char a = trans_char(transA), b = trans_char(transB);
  xkblas_zgemm_native_(&a, &b, &M, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_zgemmt_native_(const char* uplo, const char* ta, const char* tb, const int* n, const int* k,
                           const Complex64_t* alpha, const Complex64_t* A, const int* lda, const Complex64_t* B,
                           const int* ldb, const Complex64_t* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const char*, const int*, const int*, const Complex64_t*,
                      const Complex64_t*, const int*, const Complex64_t*, const int*, const Complex64_t*, Complex64_t*,
                const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zgemmt_");
  f(uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zgemmt_native(int uplo, int transA, int transB, int N, int K, const Complex64_t* alpha,
                         const Complex64_t* A, int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta,
                         Complex64_t* C, int LDC)
{
  char u = uplo_char(uplo), a = trans_char(transA), b = trans_char(transB);
  xkblas_zgemmt_native_(&u, &a, &b, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_ztrsm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B,
                          const int* ldb)
{
  typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const Complex64_t*,
                      const Complex64_t*, const int*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "ztrsm_");
  f(side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}

int xkblas_ztrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                        const Complex64_t* A, int LDA, Complex64_t* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), a = trans_char(transA),
                                                        d = diag == CblasUnit ? 'U' : 'N';
  xkblas_ztrsm_native_(&s, &u, &a, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

void xkblas_ztrmm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex64_t* alpha, const Complex64_t* A, const int* lda, Complex64_t* B,
                          const int* ldb)
{
  typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const Complex64_t*,
                      const Complex64_t*, const int*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "ztrmm_");
  f(side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}

int xkblas_ztrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                        const Complex64_t* A, int LDA, Complex64_t* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), a = trans_char(transA),
                                                        d = diag == CblasUnit ? 'U' : 'N';
  xkblas_ztrmm_native_(&s, &u, &a, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

void xkblas_zsymm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                          const Complex64_t* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex64_t*, const Complex64_t*,
                      const int*, const Complex64_t*, const int*, const Complex64_t*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zsymm_");
  f(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zsymm_native(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo);
  xkblas_zsymm_native_(&s, &u, &M, &N, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_zhemm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                          const Complex64_t* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex64_t*, const Complex64_t*,
                      const int*, const Complex64_t*, const int*, const Complex64_t*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zhemm_");
  f(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zhemm_native(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo);
  xkblas_zhemm_native_(&s, &u, &M, &N, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_zsyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                          const Complex64_t* A, const int* lda, const Complex64_t* beta, Complex64_t* C,
                          const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex64_t*, const Complex64_t*,
                      const int*, const Complex64_t*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zsyrk_");
  f(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

int xkblas_zsyrk_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                        const Complex64_t* beta, Complex64_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_zsyrk_native_(&u, &t, &N, &K, alpha, A, &LDA, beta, C, &LDC);
  return 0;
}

void xkblas_zherk_native_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
                          const Complex64_t* A, const int* lda, const double* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const double*, const Complex64_t*, const int*,
                      const double*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zherk_");
  f(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

int xkblas_zherk_native(int uplo, int trans, int N, int K, const double* alpha, const Complex64_t* A, int LDA,
                        const double* beta, Complex64_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_zherk_native_(&u, &t, &N, &K, alpha, A, &LDA, beta, C, &LDC);
  return 0;
}

void xkblas_zsyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                           const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                           const Complex64_t* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex64_t*, const Complex64_t*,
                      const int*, const Complex64_t*, const int*, const Complex64_t*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zsyr2k_");
  f(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zsyr2k_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                         const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_zsyr2k_native_(&u, &t, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_zher2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex64_t* alpha,
                           const Complex64_t* A, const int* lda, const Complex64_t* B, const int* ldb,
                           const double* beta, Complex64_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex64_t*, const Complex64_t*,
                      const int*, const Complex64_t*, const int*, const double*, Complex64_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zher2k_");
  f(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_zher2k_native(int uplo, int trans, int N, int K, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                         const Complex64_t* B, int LDB, const double* beta, Complex64_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_zher2k_native_(&u, &t, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_zpotrf_native_(const char* uplo, const int* n, Complex64_t* A, const int* lda, int* info)
{
  typedef void (*f_t)(const char*, const int*, Complex64_t*, const int*, int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "zpotrf_");
  f(uplo, n, A, lda, info);
}

int xkblas_zpotrf_native(int uplo, int N, Complex64_t* A, int LDA)
{
  char u = uplo_char(uplo); int info = 0;
  xkblas_zpotrf_native_(&u, &N, A, &LDA, &info);
  return info;
}

void xkblas_cgemm_native_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                          const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B,
                          const int* ldb, const Complex32_t* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const int*, const Complex32_t*,
                      const Complex32_t*, const int*, const Complex32_t*, const int*, const Complex32_t*, Complex32_t*,
                const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "cgemm_");
  f(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_cgemm_native(int transA, int transB, int M, int N, int K, const Complex32_t* alpha, const Complex32_t* A,
                        int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  char a = trans_char(transA), b = trans_char(transB);
  xkblas_cgemm_native_(&a, &b, &M, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_cgemmt_native_(const char* uplo, const char* ta, const char* tb, const int* n, const int* k,
                           const Complex32_t* alpha, const Complex32_t* A, const int* lda, const Complex32_t* B,
                           const int* ldb, const Complex32_t* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const char*, const int*, const int*, const Complex32_t*,
                      const Complex32_t*, const int*, const Complex32_t*, const int*, const Complex32_t*, Complex32_t*,
                const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "cgemmt_");
  f(uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_cgemmt_native(int uplo, int transA, int transB, int N, int K, const Complex32_t* alpha,
                         const Complex32_t* A, int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta,
                         Complex32_t* C, int LDC)
{
  char u = uplo_char(uplo), a = trans_char(transA), b = trans_char(transB);
  xkblas_cgemmt_native_(&u, &a, &b, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_ctrsm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B,
                          const int* ldb)
{
  typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const Complex32_t*,
                      const Complex32_t*, const int*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "ctrsm_");
  f(side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}

int xkblas_ctrsm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                        const Complex32_t* A, int LDA, Complex32_t* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), a = trans_char(transA),
                                                        d = diag == CblasUnit ? 'U' : 'N';
  xkblas_ctrsm_native_(&s, &u, &a, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

void xkblas_ctrmm_native_(const char* side, const char* uplo, const char* ta, const char* diag, const int* m,
                          const int* n, const Complex32_t* alpha, const Complex32_t* A, const int* lda, Complex32_t* B,
                          const int* ldb)
{
  typedef void (*f_t)(const char*, const char*, const char*, const char*, const int*, const int*, const Complex32_t*,
                      const Complex32_t*, const int*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "ctrmm_");
  f(side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}

int xkblas_ctrmm_native(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                        const Complex32_t* A, int LDA, Complex32_t* B, int LDB)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo), a = trans_char(transA),
                                                        d = diag == CblasUnit ? 'U' : 'N';
  xkblas_ctrmm_native_(&s, &u, &a, &d, &N, &NRHS, alpha, A, &LDA, B, &LDB);
  return 0;
}

void xkblas_csymm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                          const Complex32_t* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex32_t*, const Complex32_t*,
                      const int*, const Complex32_t*, const int*, const Complex32_t*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "csymm_");
  f(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_csymm_native(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo);
  xkblas_csymm_native_(&s, &u, &M, &N, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_chemm_native_(const char* side, const char* uplo, const int* m, const int* n, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                          const Complex32_t* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex32_t*, const Complex32_t*,
                      const int*, const Complex32_t*, const int*, const Complex32_t*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "chemm_");
  f(side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_chemm_native(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  char s = side == CblasLeft ? 'L' : 'R', u = uplo_char(uplo);
  xkblas_chemm_native_(&s, &u, &M, &N, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_csyrk_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                          const Complex32_t* A, const int* lda, const Complex32_t* beta, Complex32_t* C,
                          const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex32_t*, const Complex32_t*,
                      const int*, const Complex32_t*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "csyrk_");
  f(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

int xkblas_csyrk_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                        const Complex32_t* beta, Complex32_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_csyrk_native_(&u, &t, &N, &K, alpha, A, &LDA, beta, C, &LDC);
  return 0;
}

void xkblas_cherk_native_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha,
                          const Complex32_t* A, const int* lda, const float* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const float*, const Complex32_t*, const int*,
                      const float*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "cherk_");
  f(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

int xkblas_cherk_native(int uplo, int trans, int N, int K, const float* alpha, const Complex32_t* A, int LDA,
                        const float* beta, Complex32_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_cherk_native_(&u, &t, &N, &K, alpha, A, &LDA, beta, C, &LDC);
  return 0;
}

void xkblas_csyr2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                           const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                           const Complex32_t* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex32_t*, const Complex32_t*,
                      const int*, const Complex32_t*, const int*, const Complex32_t*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "csyr2k_");
  f(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_csyr2k_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                         const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_csyr2k_native_(&u, &t, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_cher2k_native_(const char* uplo, const char* trans, const int* n, const int* k, const Complex32_t* alpha,
                           const Complex32_t* A, const int* lda, const Complex32_t* B, const int* ldb,
                           const float* beta, Complex32_t* C, const int* ldc)
{
  typedef void (*f_t)(const char*, const char*, const int*, const int*, const Complex32_t*, const Complex32_t*,
                      const int*, const Complex32_t*, const int*, const float*, Complex32_t*, const int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "cher2k_");
  f(uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int xkblas_cher2k_native(int uplo, int trans, int N, int K, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                         const Complex32_t* B, int LDB, const float* beta, Complex32_t* C, int LDC)
{
  char u = uplo_char(uplo), t = trans_char(trans);
  xkblas_cher2k_native_(&u, &t, &N, &K, alpha, A, &LDA, B, &LDB, beta, C, &LDC);
  return 0;
}

void xkblas_cpotrf_native_(const char* uplo, const int* n, Complex32_t* A, const int* lda, int* info)
{
  typedef void (*f_t)(const char*, const int*, Complex32_t*, const int*, int*);
  static f_t f = nullptr;
  if (!f) xkblas_load_sym((void**)&f, "cpotrf_");
  f(uplo, n, A, lda, info);
}

int xkblas_cpotrf_native(int uplo, int N, Complex32_t* A, int LDA)
{
  char u = uplo_char(uplo); int info = 0;
  xkblas_cpotrf_native_(&u, &N, A, &LDA, &info);
  return info;
}

}  // extern "C"
