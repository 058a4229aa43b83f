/*
 * oracle.c — CPU restatement of XKBlas's tiled BLAS-3 path.  TEST
 * INFRASTRUCTURE ONLY: it is the checker for libxkblas_amd.so and is never
 * linked into, loaded by or called from the product.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it.
 *
 * What it restates, with the reference file:line each part follows:
 *   - tile bodies: the reference's tile tasks call a vendor BLAS on one tile
 *     (blas/ztask_gemm.c:161-173, ztask_syrk.c:113-124, ztask_trsm.c:144-159).
 *     Here they are textbook loops (gemm, syrk, trsm in all 8 cases,
 *     unblocked Cholesky);
 *   - tile algorithms: the loop nests of blas/zgemm.c:279-367,
 *     blas/zsyrk.c:206-317, blas/ztrsm.c:240-496 and the (unbuilt) right-
 *     looking Cholesky of blas/zpotrf.c:89-187;
 *   - flop model: blas/flops.h:52-95;
 *   - symmetric generator: testing/zplgsy.c:37-99 (bit-exact, including the
 *     single-precision multiplier and the per-column restart);
 *   - host BLAS loader: dlopen + dlsym of the Fortran symbols as in
 *     blas/xkblas.c:2678-2695 (LP64 integers);
 *   - double-double GEMM (TwoSum / TwoProd) as an oracle above FP64, for the
 *     accuracy statements of the split-precision modes.
 *
 * Pinning (each is checked by tests/test_oracle.py):
 *   - tile bodies and tiled drivers vs the vendor host BLAS (MKL) within the
 *     reference tests' residual bounds, and bit-exact on small-integer inputs
 *     whose products and sums are exact in FP64;
 *   - tiled Cholesky vs LAPACK dpotrf and on A = L0 L0^T with a known integer
 *     factor (exact);
 *   - double-double GEMM vs exact integer arithmetic (Python ints) on inputs
 *     whose exact products need more than 53 bits;
 *   - zplgsy vs a Python big-integer restatement of the same LCG.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { NoTrans = 111, Trans = 112, ConjTrans = 113 };
enum { Upper = 121, Lower = 122 };
enum { NonUnit = 131, Unit = 132 };
enum { Left = 141, Right = 142 };

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ---------------------------------------------------------------- flops */
/* blas/flops.h:52-53,194 */
double oracle_flops_gemm(double m, double n, double k) { return 2.0 * m * n * k; }
/* blas/flops.h:60-61,210: k*n*(n+1) */
double oracle_flops_syrk(double k, double n) { return k * n * (n + 1.0); }
/* Textbook TRSM flops m^2 n (Left) / m n^2 (Right).  The reference's macro
 * aliases FADDS_TRSM to FMULS_TRMM (blas/flops.h:77-78) and so counts
 * n*m*(m+1) for the Left case; oracle_flops_trsm_ref returns that figure. */
double oracle_flops_trsm(int side, double m, double n) { return side == Left ? m * m * n : m * n * n; }
double oracle_flops_trsm_ref(int side, double m, double n)
{
  return side == Left ? n * m * (m + 1.0) : m * n * (n + 1.0);
}
/* blas/flops.h:94-95,251: n^3/3 + n^2/2 + n/6 */
double oracle_flops_potrf(double n)
{
  double fm = n * (((1. / 6.) * n + 0.5) * n + (1. / 3.));
  double fa = n * (((1. / 6.) * n) * n - (1. / 6.));
  return fm + fa;
}

/* ---------------------------------------------------------- tile bodies */
static inline double opval(const double* A, int lda, int trans, int i, int j)
{
  return trans == NoTrans ? A[IDX(i, j, lda)] : A[IDX(j, i, lda)];
}

/* C <- alpha op(A) op(B) + beta C on one tile (BLAS semantics: beta == 0
 * overwrites C, alpha == 0 skips the product). */
void oracle_tile_dgemm(int transA, int transB, int m, int n, int k, double alpha,
                       const double* A, int lda, const double* B, int ldb,
                       double beta, double* C, int ldc)
{
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      double c = (beta == 0.0) ? 0.0 : beta * C[IDX(i, j, ldc)];
      C[IDX(i, j, ldc)] = c;
    }
    if (alpha == 0.0) continue;
    for (int l = 0; l < k; ++l) {
      double b = alpha * opval(B, ldb, transB, l, j);
      if (transA == NoTrans) {
        const double* a = A + IDX(0, l, lda);
        double* c = C + IDX(0, j, ldc);
        for (int i = 0; i < m; ++i) c[i] += a[i] * b;
      } else {
        for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] += A[IDX(l, i, lda)] * b;
      }
    }
  }
}

/* C <- alpha op(A) op(A)^T + beta C on the uplo triangle of an n x n tile. */
void oracle_tile_dsyrk(int uplo, int trans, int n, int k, double alpha,
                       const double* A, int lda, double beta, double* C, int ldc)
{
  for (int j = 0; j < n; ++j) {
    int i0 = uplo == Lower ? j : 0;
    int i1 = uplo == Lower ? n : j + 1;
    for (int i = i0; i < i1; ++i) {
      double s = 0.0;
      for (int l = 0; l < k; ++l) {
        double ail = trans == NoTrans ? A[IDX(i, l, lda)] : A[IDX(l, i, lda)];
        double ajl = trans == NoTrans ? A[IDX(j, l, lda)] : A[IDX(l, j, lda)];
        s += ail * ajl;
      }
      double c = (beta == 0.0) ? 0.0 : beta * C[IDX(i, j, ldc)];
      C[IDX(i, j, ldc)] = (alpha == 0.0) ? c : c + alpha * s;
    }
  }
}

/* Triangular solve on one tile: op(A) X = alpha B (Left) or X op(A) = alpha B
 * (Right); B is m x n and is overwritten by X. */
void oracle_tile_dtrsm(int side, int uplo, int transA, int diag, int m, int n,
                       double alpha, const double* A, int lda, double* B, int ldb)
{
  int tr = transA != NoTrans;
  /* The effective operator T = op(A) is lower triangular iff uplo==Lower
   * xor transposed. */
  int lower_eff = (uplo == Lower) != tr;
  if (side == Left) {
    for (int j = 0; j < n; ++j) {
      double* b = B + IDX(0, j, ldb);
      for (int i = 0; i < m; ++i) b[i] *= alpha;
      if (lower_eff) {
        for (int i = 0; i < m; ++i) {
          double s = b[i];
          for (int l = 0; l < i; ++l) s -= opval(A, lda, transA, i, l) * b[l];
          b[i] = diag == Unit ? s : s / opval(A, lda, transA, i, i);
        }
      } else {
        for (int i = m - 1; i >= 0; --i) {
          double s = b[i];
          for (int l = i + 1; l < m; ++l) s -= opval(A, lda, transA, i, l) * b[l];
          b[i] = diag == Unit ? s : s / opval(A, lda, transA, i, i);
        }
      }
    }
  } else {
    /* X T = alpha B: solve column by column of X.  X(:,j) T(j,j) =
     * alpha B(:,j) - sum_{l != j, T(l,j) != 0} X(:,l) T(l,j). */
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) B[IDX(i, j, ldb)] *= alpha;
    if (lower_eff) { /* T(l,j) != 0 for l >= j: go from the last column */
      for (int j = n - 1; j >= 0; --j) {
        for (int l = j + 1; l < n; ++l) {
          double t = opval(A, lda, transA, l, j);
          for (int i = 0; i < m; ++i) B[IDX(i, j, ldb)] -= B[IDX(i, l, ldb)] * t;
        }
        if (diag != Unit) {
          double d = opval(A, lda, transA, j, j);
          for (int i = 0; i < m; ++i) B[IDX(i, j, ldb)] /= d;
        }
      }
    } else {
      for (int j = 0; j < n; ++j) {
        for (int l = 0; l < j; ++l) {
          double t = opval(A, lda, transA, l, j);
          for (int i = 0; i < m; ++i) B[IDX(i, j, ldb)] -= B[IDX(i, l, ldb)] * t;
        }
        if (diag != Unit) {
          double d = opval(A, lda, transA, j, j);
          for (int i = 0; i < m; ++i) B[IDX(i, j, ldb)] /= d;
        }
      }
    }
  }
}

/* Unblocked Cholesky of one tile (LAPACK dpotf2 semantics).  Returns 0 or
 * the 1-based order of the first non-positive leading minor. */
int oracle_tile_dpotrf(int uplo, int n, double* A, int lda)
{
  for (int j = 0; j < n; ++j) {
    double d = A[IDX(j, j, lda)];
    for (int l = 0; l < j; ++l) {
      double v = uplo == Lower ? A[IDX(j, l, lda)] : A[IDX(l, j, lda)];
      d -= v * v;
    }
    if (!(d > 0.0)) {
      A[IDX(j, j, lda)] = d;
      return j + 1;
    }
    d = sqrt(d);
    A[IDX(j, j, lda)] = d;
    for (int i = j + 1; i < n; ++i) {
      if (uplo == Lower) {
        double s = A[IDX(i, j, lda)];
        for (int l = 0; l < j; ++l) s -= A[IDX(i, l, lda)] * A[IDX(j, l, lda)];
        A[IDX(i, j, lda)] = s / d;
      } else {
        double s = A[IDX(j, i, lda)];
        for (int l = 0; l < j; ++l) s -= A[IDX(l, i, lda)] * A[IDX(l, j, lda)];
        A[IDX(j, i, lda)] = s / d;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------- tiled drivers */
static inline int tsz(int i, int nt, int total, int nb) { return i == nt - 1 ? total - i * nb : nb; }
static inline int ntiles(int total, int nb) { return (total + nb - 1) / nb; }

/* Tile GEMM (blas/zgemm.c:279-367): one tile task per (m,n,k), beta applied at
 * k == 0 only.  Tile (i,j) of a matrix starts at X + i*NB + j*NB*LD
 * (blas/xkblas.c:490). */
int oracle_dgemm_tiled(int transA, int transB, int M, int N, int K, double alpha,
                       const double* A, int LDA, const double* B, int LDB,
                       double beta, double* C, int LDC, int NB)
{
  if (M == 0 || N == 0 || ((alpha == 0.0 || K == 0) && beta == 1.0)) return 0;
  int mt = ntiles(M, NB), nt = ntiles(N, NB), kt = ntiles(K, NB);
  if (K == 0) { /* no k tile: only the beta scaling remains */
    oracle_tile_dgemm(transA, transB, M, N, 0, alpha, A, LDA, B, LDB, beta, C, LDC);
    return 0;
  }
  for (int m = 0; m < mt; ++m) {
    int mm = tsz(m, mt, M, NB);
    for (int n = 0; n < nt; ++n) {
      int nn = tsz(n, nt, N, NB);
      for (int k = 0; k < kt; ++k) {
        int kk = tsz(k, kt, K, NB);
        double b = k == 0 ? beta : 1.0;
        const double* Ak = transA == NoTrans ? A + IDX(m * NB, k * NB, LDA) : A + IDX(k * NB, m * NB, LDA);
        const double* Bk = transB == NoTrans ? B + IDX(k * NB, n * NB, LDB) : B + IDX(n * NB, k * NB, LDB);
        oracle_tile_dgemm(transA, transB, mm, nn, kk, alpha, Ak, LDA, Bk, LDB, b, C + IDX(m * NB, n * NB, LDC), LDC);
      }
    }
  }
  return 0;
}

/* Tile SYRK (blas/zsyrk.c:206-317): diagonal tiles are SYRK tasks, the other
 * tiles of the uplo triangle are GEMM tasks; beta at k == 0 only. */
int oracle_dsyrk_tiled(int uplo, int trans, int N, int K, double alpha,
                       const double* A, int LDA, double beta, double* C, int LDC, int NB)
{
  if (N == 0 || ((alpha == 0.0 || K == 0) && beta == 1.0)) return 0;
  int nt = ntiles(N, NB), kt = ntiles(K, NB);
  if (K == 0) {
    oracle_tile_dsyrk(uplo, trans, N, 0, alpha, A, LDA, beta, C, LDC);
    return 0;
  }
  /* Tile (i,k) of op(A): rows of op(A) are rows of A (NoTrans) or columns. */
#define OPA(i, k) (trans == NoTrans ? A + IDX((i) * NB, (k) * NB, LDA) : A + IDX((k) * NB, (i) * NB, LDA))
  for (int n = 0; n < nt; ++n) {
    int nn = tsz(n, nt, N, NB);
    for (int k = 0; k < kt; ++k) {
      int kk = tsz(k, kt, K, NB);
      oracle_tile_dsyrk(uplo, trans, nn, kk, alpha, OPA(n, k), LDA, k == 0 ? beta : 1.0,
                        C + IDX(n * NB, n * NB, LDC), LDC);
    }
    for (int m = n + 1; m < nt; ++m) {
      int mm = tsz(m, nt, N, NB);
      for (int k = 0; k < kt; ++k) {
        int kk = tsz(k, kt, K, NB);
        double b = k == 0 ? beta : 1.0;
        int tA = trans == NoTrans ? NoTrans : Trans;
        int tB = trans == NoTrans ? Trans : NoTrans;
        if (uplo == Lower) /* C(m,n) += op(A)(m,k) op(A)(n,k)^T */
          oracle_tile_dgemm(tA, tB, mm, nn, kk, alpha, OPA(m, k), LDA, OPA(n, k), LDA, b,
                            C + IDX(m * NB, n * NB, LDC), LDC);
        else /* C(n,m) += op(A)(n,k) op(A)(m,k)^T */
          oracle_tile_dgemm(tA, tB, nn, mm, kk, alpha, OPA(n, k), LDA, OPA(m, k), LDA, b,
                            C + IDX(n * NB, m * NB, LDC), LDC);
      }
    }
  }
#undef OPA
  return 0;
}

/* Tile TRSM (blas/ztrsm.c:240-496).  The reference writes eight loop nests;
 * they reduce to: walk the diagonal tiles of the effective operator
 * T = op(A) in solve order, solve the diagonal tile against every tile of B
 * in that block row (Left) / column (Right), then update the remaining block
 * rows / columns with a GEMM by -1.  alpha is folded in as in the reference:
 * the first step solves with alpha and its updates scale the rest of B by
 * alpha (lalpha, blas/ztrsm.c:250,282,316,349,383,445).  The reference's two
 * Right/Trans cases instead solve every step with alpha and update with
 * -1/alpha (blas/ztrsm.c:421,479); that is algebraically the same and is not
 * reproduced.  alpha == 0 sets B to zero (BLAS), where the reference would
 * divide by zero. */
int oracle_dtrsm_tiled(int side, int uplo, int transA, int diag, int M, int N, double alpha,
                       const double* A, int LDA, double* B, int LDB, int NB)
{
  if (M == 0 || N == 0) return 0;
  if (alpha == 0.0) {
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < M; ++i) B[IDX(i, j, LDB)] = 0.0;
    return 0;
  }
  int tr = transA != NoTrans;
  int lower_eff = (uplo == Lower) != tr;
  int mt = ntiles(M, NB), nt = ntiles(N, NB);
  /* Tile (i,j) of T = op(A) is tile (i,j) of A, or tile (j,i) transposed. */
#define TT(i, j) (tr ? A + IDX((j) * NB, (i) * NB, LDA) : A + IDX((i) * NB, (j) * NB, LDA))
#define BT(i, j) (B + IDX((i) * NB, (j) * NB, LDB))
  if (side == Left) {
    int at = mt; /* T is M x M */
    for (int s = 0; s < at; ++s) {
      int k = lower_eff ? s : at - 1 - s;
      int kk = tsz(k, at, M, NB);
      double la = s == 0 ? alpha : 1.0;
      for (int n = 0; n < nt; ++n)
        oracle_tile_dtrsm(Left, uplo, transA, diag, kk, tsz(n, nt, N, NB), la, TT(k, k), LDA, BT(k, n), LDB);
      for (int r = s + 1; r < at; ++r) {
        int m = lower_eff ? r : at - 1 - r;
        for (int n = 0; n < nt; ++n)
          oracle_tile_dgemm(transA, NoTrans, tsz(m, at, M, NB), tsz(n, nt, N, NB), kk, -1.0,
                            TT(m, k), LDA, BT(k, n), LDB, la, BT(m, n), LDB);
      }
    }
  } else {
    int at = nt; /* T is N x N; X T = B walks columns */
    for (int s = 0; s < at; ++s) {
      int k = lower_eff ? at - 1 - s : s;
      int kk = tsz(k, at, N, NB);
      double la = s == 0 ? alpha : 1.0;
      for (int m = 0; m < mt; ++m)
        oracle_tile_dtrsm(Right, uplo, transA, diag, tsz(m, mt, M, NB), kk, la, TT(k, k), LDA, BT(m, k), LDB);
      for (int r = s + 1; r < at; ++r) {
        int n = lower_eff ? at - 1 - r : r;
        for (int m = 0; m < mt; ++m)
          oracle_tile_dgemm(NoTrans, transA, tsz(m, mt, M, NB), tsz(n, at, N, NB), kk, -1.0,
                            BT(m, k), LDB, TT(k, n), LDA, la, BT(m, n), LDB);
      }
    }
  }
#undef TT
#undef BT
  return 0;
}

/* Right-looking tile Cholesky (blas/zpotrf.c:89-187).  Lower:
 * POTRF(k,k); TRSM(Right,Lower,Trans,NonUnit) on tiles (m,k), m>k;
 * SYRK(Lower,NoTrans,-1) on (n,n) and GEMM(NoTrans,Trans,-1) on (m,n), m>n>k.
 * Upper is the transpose.  Returns the LAPACK info (0 or first bad order). */
int oracle_dpotrf_tiled(int uplo, int N, double* A, int LDA, int NB)
{
  int nt = ntiles(N, NB);
#define AT(i, j) (A + IDX((i) * NB, (j) * NB, LDA))
  for (int k = 0; k < nt; ++k) {
    int kk = tsz(k, nt, N, NB);
    int info = oracle_tile_dpotrf(uplo, kk, AT(k, k), LDA);
    if (info) return k * NB + info;
    for (int m = k + 1; m < nt; ++m) {
      int mm = tsz(m, nt, N, NB);
      if (uplo == Lower)
        oracle_tile_dtrsm(Right, Lower, Trans, NonUnit, mm, NB, 1.0, AT(k, k), LDA, AT(m, k), LDA);
      else
        oracle_tile_dtrsm(Left, Upper, Trans, NonUnit, NB, mm, 1.0, AT(k, k), LDA, AT(k, m), LDA);
    }
    for (int n = k + 1; n < nt; ++n) {
      int nn = tsz(n, nt, N, NB);
      if (uplo == Lower)
        oracle_tile_dsyrk(Lower, NoTrans, nn, NB, -1.0, AT(n, k), LDA, 1.0, AT(n, n), LDA);
      else
        oracle_tile_dsyrk(Upper, Trans, nn, NB, -1.0, AT(k, n), LDA, 1.0, AT(n, n), LDA);
      for (int m = n + 1; m < nt; ++m) {
        int mm = tsz(m, nt, N, NB);
        if (uplo == Lower)
          oracle_tile_dgemm(NoTrans, Trans, mm, nn, NB, -1.0, AT(m, k), LDA, AT(n, k), LDA, 1.0, AT(m, n), LDA);
        else
          oracle_tile_dgemm(Trans, NoTrans, nn, mm, NB, -1.0, AT(k, n), LDA, AT(k, m), LDA, 1.0, AT(n, m), LDA);
      }
    }
  }
#undef AT
  return 0;
}

/* --------------------------------------------------- double-double GEMM */
static inline void two_sum(double a, double b, double* s, double* e)
{
  double x = a + b, z = x - a;
  *s = x;
  *e = (a - (x - z)) + (b - z);
}
static inline void dd_add(double* hi, double* lo, double ph, double pl)
{
  double s, e;
  two_sum(*hi, ph, &s, &e);
  e += *lo + pl;
  double t = s + e;
  *lo = e - (t - s);
  *hi = t;
}

/* C_dd = alpha op(A) op(B) + beta C in double-double: every product by
 * TwoProd (fma) and every sum by TwoSum, so the result carries ~106 bits. */
void oracle_dgemm_dd(int transA, int transB, int m, int n, int k, double alpha,
                     const double* A, int lda, const double* B, int ldb, double beta,
                     const double* C, int ldc, double* Chi, double* Clo)
{
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double hi = 0.0, lo = 0.0;
      for (int l = 0; l < k; ++l) {
        double a = opval(A, lda, transA, i, l), b = opval(B, ldb, transB, l, j);
        double p = a * b, pe = fma(a, b, -p);
        dd_add(&hi, &lo, p, pe);
      }
      /* times alpha: (hi+lo)*alpha with the error term of hi*alpha */
      double ph = hi * alpha, pe = fma(hi, alpha, -ph) + lo * alpha;
      double rh = 0.0, rl = 0.0;
      dd_add(&rh, &rl, ph, pe);
      if (beta != 0.0) {
        double c = C[IDX(i, j, ldc)];
        double qh = c * beta, qe = fma(c, beta, -qh);
        dd_add(&rh, &rl, qh, qe);
      }
      Chi[IDX(i, j, m)] = rh;
      Clo[IDX(i, j, m)] = rl;
    }
}

/* ------------------------------------------------------------ generator */
/* testing/zplgsy.c:37-99: LCG a=6364136223846793005, c=1, jumped by N for
 * every column (the reference restarts the stream at the same state for
 * each column), value 0.5f - ran * 5.4210108624275222e-20f computed in
 * single precision, lower part generated then mirrored, bump on diagonal. */
static unsigned long long lcg_jump(unsigned long long n, unsigned long long seed)
{
  unsigned long long a = 6364136223846793005ULL, c = 1ULL, ran = seed;
  for (; n; n >>= 1) {
    if (n & 1) ran = a * ran + c;
    c *= (a + 1);
    a *= a;
  }
  return ran;
}

void oracle_plgsy(double bump, int N, double* A, int lda, unsigned long long seed)
{
  for (int j = 0; j < N; ++j) {
    unsigned long long ran = lcg_jump((unsigned long long)N, seed);
    for (int i = j; i < N; ++i) {
      float v = 0.5f - (float)ran * 5.4210108624275222e-20f;
      A[IDX(i, j, lda)] = v;
      ran = 6364136223846793005ULL * ran + 1ULL;
    }
  }
  for (int j = 0; j < N; ++j) {
    A[IDX(j, j, lda)] += bump;
    for (int i = 0; i < j; ++i) A[IDX(i, j, lda)] = A[IDX(j, i, lda)];
  }
}

/* ------------------------------------------------------- host BLAS path */
typedef void (*dgemm_f)(const char*, const char*, const int*, const int*, const int*, const double*,
                        const double*, const int*, const double*, const int*, const double*, double*,
                        const int*);
typedef void (*dsyrk_f)(const char*, const char*, const int*, const int*, const double*, const double*,
                        const int*, const double*, double*, const int*);
typedef void (*dtrsm_f)(const char*, const char*, const char*, const char*, const int*, const int*,
                        const double*, const double*, const int*, double*, const int*);
typedef void (*dpotrf_f)(const char*, const int*, double*, const int*, int*);

static void* g_blas = 0;
static dgemm_f p_dgemm = 0;
static dsyrk_f p_dsyrk = 0;
static dtrsm_f p_dtrsm = 0;
static dpotrf_f p_dpotrf = 0;

/* blas/xkblas.c:2678-2695: dlopen the library named by XKBLAS_BLASLIB and
 * resolve the Fortran symbols.  Returns 0 on success. */
int oracle_hostblas_open(const char* path)
{
  if (g_blas) return 0;
  g_blas = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!g_blas) return -1;
  p_dgemm = (dgemm_f)dlsym(g_blas, "dgemm_");
  p_dsyrk = (dsyrk_f)dlsym(g_blas, "dsyrk_");
  p_dtrsm = (dtrsm_f)dlsym(g_blas, "dtrsm_");
  p_dpotrf = (dpotrf_f)dlsym(g_blas, "dpotrf_");
  if (!p_dgemm || !p_dsyrk || !p_dtrsm || !p_dpotrf) {
    dlclose(g_blas);
    g_blas = 0;
    return -2;
  }
  return 0;
}

static char tr_char(int t) { return t == NoTrans ? 'N' : (t == Trans ? 'T' : 'C'); }
static char ul_char(int u) { return u == Lower ? 'L' : 'U'; }

int oracle_host_dgemm(int transA, int transB, int m, int n, int k, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, double* C, int ldc)
{
  if (!p_dgemm) return -1;
  char ta = tr_char(transA), tb = tr_char(transB);
  p_dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
  return 0;
}

int oracle_host_dsyrk(int uplo, int trans, int n, int k, double alpha, const double* A, int lda,
                      double beta, double* C, int ldc)
{
  if (!p_dsyrk) return -1;
  char u = ul_char(uplo), t = tr_char(trans);
  p_dsyrk(&u, &t, &n, &k, &alpha, A, &lda, &beta, C, &ldc);
  return 0;
}

int oracle_host_dtrsm(int side, int uplo, int transA, int diag, int m, int n, double alpha,
                      const double* A, int lda, double* B, int ldb)
{
  if (!p_dtrsm) return -1;
  char s = side == Left ? 'L' : 'R', u = ul_char(uplo), t = tr_char(transA), d = diag == Unit ? 'U' : 'N';
  p_dtrsm(&s, &u, &t, &d, &m, &n, &alpha, A, &lda, B, &ldb);
  return 0;
}

int oracle_host_dpotrf(int uplo, int n, double* A, int lda)
{
  if (!p_dpotrf) return -1000000;
  char u = ul_char(uplo);
  int info = 0;
  p_dpotrf(&u, &n, A, &lda, &info);
  return info;
}

/* Double-double value of selected entries (rows[e], cols[e]) of
 * alpha op(A) op(B) + beta C0, for accuracy statements on large products
 * where the full double-double GEMM would be too slow. */
void oracle_dd_entries(int transA, int transB, int k, double alpha, const double* A, int lda, const double* B,
                       int ldb, double beta, const double* C0, int ldc, int ne, const int* rows, const int* cols,
                       double* hi_out, double* lo_out)
{
  for (int e = 0; e < ne; ++e) {
    const int i = rows[e], j = cols[e];
    double hi = 0.0, lo = 0.0;
    for (int l = 0; l < k; ++l) {
      double a = opval(A, lda, transA, i, l), b = opval(B, ldb, transB, l, j);
      double p = a * b, pe = fma(a, b, -p);
      dd_add(&hi, &lo, p, pe);
    }
    double ph = hi * alpha, pe = fma(hi, alpha, -ph) + lo * alpha;
    double rh = 0.0, rl = 0.0;
    dd_add(&rh, &rl, ph, pe);
    if (beta != 0.0) {
      double c = C0[IDX(i, j, ldc)];
      double qh = c * beta, qe = fma(c, beta, -qh);
      dd_add(&rh, &rl, qh, qe);
    }
    hi_out[e] = rh;
    lo_out[e] = rl;
  }
}
