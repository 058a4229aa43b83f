// blas3.cpp — BLAS-3 algorithms on device-resident matrices (host side).
//
// GEMM / SYRK (reference blas/zgemm.c:279-367, zsyrk.c:206-317): the
// reference inserts one task per (m,n,k) tile with beta at k == 0; here one
// launch covers every output tile and runs the whole k loop in registers, so
// beta is applied once, exactly as the tile sequence does.  SYRK is the same
// kernel restricted to the uplo triangle of C (diagonal tiles masked).
//
// TRSM (reference blas/ztrsm.c:240-496): the reference alternates a diagonal
// tile solve with GEMM updates by -1, folding alpha into the first step.  The
// same sweep is kept, but each diagonal solve is a GEMM with the explicit
// inverse of the diagonal block (all inverses computed up front by a
// recursive TRTRI whose leaves are 64 x 64 single-wave kernels), so every
// step runs on MFMA instead of a latency-bound substitution.
//
// POTRF (reference blas/zpotrf.c:89-187, unbuilt there): right-looking
// blocked Cholesky with block order nb.  Step k factors the diagonal block
// (recursive: leaves of order 64, inverse built alongside), forms the panel
// as a GEMM with the inverse, and applies the rank-nb update to the trailing
// triangle in one launch.  With lookahead the next diagonal block is updated
// first and factored on a side stream while the bulk update runs.
#include "blas3.h"
#include "kernels.h"
#include "split.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace xk {

template <typename T>
void dev_gemm(const Exec& x, int transA, int transB, int m, int n, int k, T alpha, const T* A, long lda,
              const T* B, long ldb, T beta, T* C, long ldc)
{
  if (m <= 0 || n <= 0) return;
  if (split_enabled<T>(x) && k > 0 && alpha != T(0)) {
    split_gemm<T>(x, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, TRI_FULL);
    return;
  }
  gemm<T>(x.s, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename T>
void dev_syrk(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, T beta, T* C,
              long ldc)
{
  if (n <= 0) return;
  const int ta = trans == NoTrans ? NoTrans : Trans;
  const int tb = trans == NoTrans ? Trans : NoTrans;
  const int tri = uplo == Lower ? TRI_LOWER : TRI_UPPER;
  if (split_enabled<T>(x) && k > 0 && alpha != T(0)) {
    split_gemm<T>(x, ta, tb, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, tri);
    return;
  }
  gemm<T>(x.s, ta, tb, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, tri);
}

// Split point of a recursive step: a multiple of 64 near n/2.
static inline int half64(int n)
{
  int h = ((n / 2) + 63) / 64 * 64;
  return h >= n ? n - 64 : h;
}

template <typename T>
static void fill_zero(hipStream_t s, int m, int n, T* Y, long ldy)
{
  // zero a block via the masked GEMM path: alpha = 0, beta = 0, k = 0
  gemm<T>(s, NoTrans, NoTrans, m, n, 0, T(0), Y, ldy, Y, ldy, T(0), Y, ldy);
}

template <typename T>
void dev_trtri(const Exec& x, hipStream_t s, int uplo, int diag, int n, const T* A, long lda, T* Tinv, long ldinv,
               int wslot)
{
  if (n <= 64) {
    trtri_leaf<T>(s, uplo, diag, n, A, lda, 0, Tinv, ldinv, 0, 1);
    return;
  }
  const int n1 = half64(n), n2 = n - n1;
  const T* A11 = A;
  const T* A22 = A + n1 + (long)n1 * lda;
  T* I11 = Tinv;
  T* I22 = Tinv + n1 + (long)n1 * ldinv;
  dev_trtri<T>(x, s, uplo, diag, n1, A11, lda, I11, ldinv, wslot);
  dev_trtri<T>(x, s, uplo, diag, n2, A22, lda, I22, ldinv, wslot);
  T* W = (T*)scratch(wslot, sizeof(T) * (size_t)n1 * n2);
  if (uplo == Lower) {
    // I21 = -I22 A21 I11
    const T* A21 = A + n1;
    T* I21 = Tinv + n1;
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n1, T(1), A21, lda, I11, ldinv, T(0), W, n2);
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n2, T(-1), I22, ldinv, W, n2, T(0), I21, ldinv);
    fill_zero<T>(s, n1, n2, Tinv + (long)n1 * ldinv, ldinv);
  } else {
    // I12 = -I11 A12 I22
    const T* A12 = A + (long)n1 * lda;
    T* I12 = Tinv + (long)n1 * ldinv;
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n2, T(1), A12, lda, I22, ldinv, T(0), W, n1);
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n1, T(-1), I11, ldinv, W, n1, T(0), I12, ldinv);
    fill_zero<T>(s, n2, n1, Tinv + n1, ldinv);
  }
}

// Recursive Cholesky of one diagonal block with its inverse (Tinv = inv(L)
// for Lower, inv(U) for Upper).
// pair != 0: A is the real form of a complex block (orders 2i, 2i + 1 stand
// for complex order i), and a failing pivot reports its complex order.
template <typename T>
static void potrf_diag(hipStream_t s, int uplo, int n, T* A, long lda, T* Tinv, long ldinv, int* info, int base,
                       int wslot, bool zeroed = false, int pair = 0)
{
  if (n <= 64) {
    potrf_leaf<T>(s, uplo, n, A, lda, 0, Tinv, ldinv, 0, 1, info, base, pair);
    return;
  }
  const int n1 = half64(n), n2 = n - n1;
  T* A11 = A;
  T* A22 = A + n1 + (long)n1 * lda;
  T* I11 = Tinv;
  T* I22 = Tinv + n1 + (long)n1 * ldinv;
  potrf_diag<T>(s, uplo, n1, A11, lda, I11, ldinv, info, base, wslot, zeroed, pair);
  T* W = (T*)scratch(wslot, sizeof(T) * (size_t)n1 * n2);
  if (uplo == Lower) {
    T* A21 = A + n1;
    T* I21 = Tinv + n1;
    copy2d<T>(s, n2, n1, A21, lda, W, n2);
    gemm<T>(s, NoTrans, Trans, n2, n1, n1, T(1), W, n2, I11, ldinv, T(0), A21, lda);   // L21 = A21 inv(L11)^T
    gemm<T>(s, NoTrans, Trans, n2, n2, n1, T(-1), A21, lda, A21, lda, T(1), A22, lda, TRI_LOWER);
    potrf_diag<T>(s, uplo, n2, A22, lda, I22, ldinv, info, base + n1, wslot, zeroed, pair);
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n1, T(1), A21, lda, I11, ldinv, T(0), W, n2);  // L21 inv(L11)
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n2, T(-1), I22, ldinv, W, n2, T(0), I21, ldinv);
    if (!zeroed) fill_zero<T>(s, n1, n2, Tinv + (long)n1 * ldinv, ldinv);
  } else {
    T* A12 = A + (long)n1 * lda;
    T* I12 = Tinv + (long)n1 * ldinv;
    copy2d<T>(s, n1, n2, A12, lda, W, n1);
    gemm<T>(s, Trans, NoTrans, n1, n2, n1, T(1), I11, ldinv, W, n1, T(0), A12, lda);     // U12 = inv(U11)^T A12
    gemm<T>(s, Trans, NoTrans, n2, n2, n1, T(-1), A12, lda, A12, lda, T(1), A22, lda, TRI_UPPER);
    potrf_diag<T>(s, uplo, n2, A22, lda, I22, ldinv, info, base + n1, wslot, zeroed, pair);
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n2, T(1), A12, lda, I22, ldinv, T(0), W, n1);   // U12 inv(U22)
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n1, T(-1), I11, ldinv, W, n1, T(0), I12, ldinv);
    if (!zeroed) fill_zero<T>(s, n2, n1, Tinv + n1, ldinv);
  }
}

enum { WS_TRSM = 4, WS_TRSMW = 5, WS_SPLITK = 31 };

// Products with few output tiles and a long inner dimension (the inverse
// products and late updates of a TRSM sweep with few right-hand sides) are
// latency-bound: each tile walks its k slices one after the other.  Split
// the inner dimension into 2, 4 or 8 chunks computed by one batched launch
// into partial products, summed by a reduction (XKBLAS_SPLITK=0 turns it
// off).  Returns false when the shape does not qualify.
static bool splitk_enabled()
{
  static const int v = getenv("XKBLAS_SPLITK") ? atoi(getenv("XKBLAS_SPLITK")) : 1;
  return v != 0;
}

template <typename T>
static bool gemm_splitk(hipStream_t s, int ta, int tb, int m, int n, int k, T alpha, const T* A, long lda,
                        const T* B, long ldb, T beta, T* C, long ldc)
{
  if (!splitk_enabled() || m <= 0 || n <= 0 || k < 512) return false;
  const long t32 = (long)((m + 31) / 32) * ((n + 31) / 32);
  if (t32 >= 256) return false;
  int parts = 1;
  for (int p = 8; p >= 2; p /= 2)
    if (k % (16 * p) == 0 && t32 * p <= 1024) {
      parts = p;
      break;
    }
  if (parts == 1) return false;
  const int kc = k / parts;
  T* W = (T*)scratch(WS_SPLITK, sizeof(T) * (size_t)parts * m * n);
  const long sA = ta != NoTrans ? (long)kc : (long)kc * lda;
  const long sB = tb != NoTrans ? (long)kc * ldb : (long)kc;
  gemm<T>(s, ta, tb, m, n, kc, alpha, A, lda, B, ldb, T(0), W, m, TRI_FULL, parts, sA, sB, (long)m * n);
  splitk_reduce<T>(s, m, n, parts, W, beta, C, ldc);
  return true;
}

void potrf_scratch_caps(int n, int nb, size_t elt, size_t cap[4])
{
  nb = std::max(64, std::min(nb, std::max(n, 1)));
  // the largest recursion node of a diagonal block of order <= nb is the top
  // one: W holds its n1 x n2 block
  const size_t h = nb > 64 ? (size_t)half64(nb) * (size_t)(nb - half64(nb)) : 0;
  cap[WS_DIAG0] = cap[WS_DIAG1] = elt * std::max<size_t>(h, 1);
  cap[WS_PANEL] = elt * std::max<size_t>((size_t)std::max(n - nb, 0) * nb, 1);
  cap[WS_INV] = elt * 2 * (size_t)nb * nb;
}

// Trailing updates of POTRF: 1 (default) paired, 0 one per panel.
// XKBLAS_POTRF_PAIR sets the initial value.
static int g_potrf_pair = getenv("XKBLAS_POTRF_PAIR") ? (atoi(getenv("XKBLAS_POTRF_PAIR")) != 0) : 1;
static int potrf_pairing() { return g_potrf_pair; }
void potrf_set_pairing(int v) { g_potrf_pair = v < 0 ? 1 : (v ? 1 : 0); }

template <typename T>
void dev_potrf(const Exec& x, int uplo, int n, T* A, long lda, int nb, int* info,
               const PotrfHooks<T>* hooks)
{
  if (n <= 0) return;
  nb = std::max(64, std::min(nb, n));
  const int nt = (n + nb - 1) / nb;
  T* const keep_inv = hooks ? hooks->keep_inv : nullptr;
  const int keep_nb = hooks ? hooks->keep_nb : 0;
  XK_CHECK(!keep_inv || (keep_nb > 0 && (nt == 1 || nb % keep_nb == 0)),
           "dev_potrf: kept inverses need a block order that divides nb");
  // inverses of two consecutive diagonal blocks (ping-pong for lookahead)
  T* inv = (T*)scratch(WS_INV, sizeof(T) * (size_t)2 * nb * nb);
  hipEvent_t ev_diag = x.ev0, ev_look = x.ev1;
  const bool own_events = !ev_diag || !ev_look;
  if (own_events) {
    XK_HIP(hipEventCreateWithFlags(&ev_diag, hipEventDisableTiming));
    XK_HIP(hipEventCreateWithFlags(&ev_look, hipEventDisableTiming));
  }
  const bool lookahead = x.s_diag != nullptr && nt > 1;
  hipStream_t sd = lookahead ? x.s_diag : x.s;
  // trailing updates of two panels at once (inner dimension 2 nb): longer
  // products, and half the split modes' CRT passes over the trailing triangle
  const bool agg = potrf_pairing() == 1;
  // split modes: the residue planes of each step's panel operand are computed
  // once for all the products that read it (its rows never change again)
  const bool split = split_enabled<T>(x);
  if (split) split_cache_clear();

  auto blk = [&](int i, int j) { return A + (long)i * nb + (long)j * nb * lda; };
  auto tsz = [&](int i) { return i == nt - 1 ? n - i * nb : nb; };
  // part 2 of a paired trailing update: the triangle of order n at C, operand U
  struct Pending {
    const T* U;
    int n, kw;
    T* C;
  };
  Pending pend{nullptr, 0, 0, nullptr};
  // columns of part 1 for a trailing triangle of order r: (1 - 1/sqrt 2) r
  // (equal flops in both parts), whole blocks, at least one block
  auto pair_split = [&](int r) {
    const int w = (int)((0.2929 * r) / nb + 0.5) * nb;
    return std::min(r, std::max(nb, w));
  };

  // factor block 0
  if (lookahead) {
    XK_HIP(hipEventRecord(ev_look, x.s));
    XK_HIP(hipStreamWaitEvent(sd, ev_look, 0));
  }
  // the chain's products issue ahead of the trailing update's waves
  if (lookahead) gemm_set_priority_stream(sd);
  // the inverses' blocks above the diagonal are never written with nonzeros:
  // one fill here replaces a fill per recursion node
  fill_zero<T>(sd, nb, 2 * nb, inv, nb);
  // the diagonal keep_nb sub-blocks of block k's inverse are the inverses of
  // the factor's diagonal blocks of order keep_nb
  auto keep = [&](int k, const T* ik) {
    if (!keep_inv) return;
    const int kk = tsz(k);
    for (int j = 0; j * keep_nb < kk; ++j) {
      const int sz = std::min(keep_nb, kk - j * keep_nb);
      const size_t g = (size_t)k * nb / keep_nb + j;
      copy2d<T>(sd, sz, sz, ik + (long)j * keep_nb * (1 + nb), nb, keep_inv + g * keep_nb * keep_nb, keep_nb);
    }
  };
  potrf_diag<T>(sd, uplo, tsz(0), blk(0, 0), lda, inv, nb, info, 0, WS_DIAG0, true);
  keep(0, inv);
  if (hooks && hooks->beside_first) hooks->beside_first();
  for (int k = 0; k < nt; ++k) {
    const int kk = tsz(k);
    T* ik = inv + (size_t)(k & 1) * nb * nb;
    if (lookahead) {
      XK_HIP(hipEventRecord(ev_diag, sd));
      XK_HIP(hipStreamWaitEvent(x.s, ev_diag, 0));
    }
    if (k == nt - 1) break;
    const int rest = n - (k + 1) * nb;
    // Paired updates: an even step applies panel k only to block column k+1
    // (the lookahead); the odd step after it applies panels k-1 and k (adjacent
    // columns of A: one operand U of inner dimension 2 nb) to block column
    // k+1 and, in two parts split by columns, to the trailing triangle: part
    // 1 (the first ~29 % of its columns, flops balanced with part 2) at once,
    // part 2 at the next step, so that every diagonal factorization on the
    // side stream has trailing work beside it on the compute stream.
    const bool pair = agg && (k & 1);
    const Pending prev = pend;
    pend.n = 0;
    const int k1 = tsz(k + 1);
    const int r2 = rest - k1;
    const int c0 = pair ? k - 1 : k;
    const int kw = (k + 1 - c0) * nb;
    T* W = (T*)scratch(WS_PANEL, sizeof(T) * (size_t)rest * kk);
    if (uplo == Lower) {
      T* P = blk(k + 1, k);  // rest x kk panel
      // the panel's rows in block row k+1 first: the lookahead reads only
      // them, so the next diagonal factorization waits for k1 rows of the
      // panel product instead of all of them
      copy2d<T>(x.s, k1, kk, P, lda, W, rest);
      dev_gemm<T>(x, NoTrans, Trans, k1, kk, kk, T(1), W, rest, ik, nb, T(0), P, lda);
      const T* U = A + (long)(k + 1) * nb + (long)c0 * nb * lda;  // rest x kw
      if (split) split_cache_register((const double*)U, lda, rest, kw, 1);
      // lookahead: update the next diagonal block and factor it on the side
      dev_syrk<T>(x, Lower, NoTrans, k1, kw, T(-1), U, lda, T(1), blk(k + 1, k + 1), lda);
      if (lookahead) XK_HIP(hipEventRecord(ev_look, x.s));
      if (r2 > 0) {
        copy2d<T>(x.s, r2, kk, P + k1, lda, W + k1, rest);
        dev_gemm<T>(x, NoTrans, Trans, r2, kk, kk, T(1), W + k1, rest, ik, nb, T(0), P + k1, lda);
      }
      // the rest is enqueued before the diagonal factorization's many small
      // launches, so the compute stream has work while the host issues them:
      // block column k+1 below its diagonal block, then the trailing triangle
      if (r2 > 0) {
        dev_gemm<T>(x, NoTrans, Trans, r2, k1, kw, T(-1), U + k1, lda, U, lda, T(1), blk(k + 2, k + 1), lda);
        T* C1 = blk(k + 2, k + 2);
        if (!agg) {
          dev_syrk<T>(x, Lower, NoTrans, r2, kw, T(-1), U + k1, lda, T(1), C1, lda);
        } else if (pair) {
          const int w1 = pair_split(r2);
          dev_syrk<T>(x, Lower, NoTrans, w1, kw, T(-1), U + k1, lda, T(1), C1, lda);
          if (w1 < r2) {
            dev_gemm<T>(x, NoTrans, Trans, r2 - w1, w1, kw, T(-1), U + k1 + w1, lda, U + k1, lda, T(1), C1 + w1, lda);
            pend = Pending{U + k1 + w1, r2 - w1, kw, C1 + w1 + (long)w1 * lda};
          }
        }
      }
      if (prev.n > 0) dev_syrk<T>(x, Lower, NoTrans, prev.n, prev.kw, T(-1), prev.U, lda, T(1), prev.C, lda);
    } else {
      T* P = blk(k, k + 1);  // kk x rest panel
      // the panel's columns in block column k+1 first (as above)
      copy2d<T>(x.s, kk, k1, P, lda, W, kk);
      dev_gemm<T>(x, Trans, NoTrans, kk, k1, kk, T(1), ik, nb, W, kk, T(0), P, lda);
      const T* U = A + (long)c0 * nb + (long)(k + 1) * nb * lda;  // kw x rest
      if (split) split_cache_register((const double*)U, lda, rest, kw, 0);
      dev_syrk<T>(x, Upper, Trans, k1, kw, T(-1), U, lda, T(1), blk(k + 1, k + 1), lda);
      if (lookahead) XK_HIP(hipEventRecord(ev_look, x.s));
      if (r2 > 0) {
        copy2d<T>(x.s, kk, r2, P + (long)k1 * lda, lda, W + (long)k1 * kk, kk);
        dev_gemm<T>(x, Trans, NoTrans, kk, r2, kk, T(1), ik, nb, W + (long)k1 * kk, kk, T(0), P + (long)k1 * lda,
                    lda);
      }
      if (r2 > 0) {
        dev_gemm<T>(x, Trans, NoTrans, k1, r2, kw, T(-1), U, lda, U + (long)k1 * lda, lda, T(1), blk(k + 1, k + 2),
                    lda);
        T* C1 = blk(k + 2, k + 2);
        if (!agg) {
          dev_syrk<T>(x, Upper, Trans, r2, kw, T(-1), U + (long)k1 * lda, lda, T(1), C1, lda);
        } else if (pair) {
          const int w1 = pair_split(r2);
          dev_syrk<T>(x, Upper, Trans, w1, kw, T(-1), U + (long)k1 * lda, lda, T(1), C1, lda);
          if (w1 < r2) {
            dev_gemm<T>(x, Trans, NoTrans, w1, r2 - w1, kw, T(-1), U + (long)k1 * lda, lda,
                        U + (long)(k1 + w1) * lda, lda, T(1), C1 + (long)w1 * lda, lda);
            pend = Pending{U + (long)(k1 + w1) * lda, r2 - w1, kw, C1 + w1 + (long)w1 * lda};
          }
        }
      }
      if (prev.n > 0) dev_syrk<T>(x, Upper, Trans, prev.n, prev.kw, T(-1), prev.U, lda, T(1), prev.C, lda);
    }
    // block column k of the factor is final (its panel and diagonal block)
    if (hooks && hooks->after_panel) hooks->after_panel(k, ik);
    if (lookahead) {
      XK_HIP(hipStreamWaitEvent(sd, ev_look, 0));
      potrf_diag<T>(sd, uplo, k1, blk(k + 1, k + 1), lda, inv + (size_t)((k + 1) & 1) * nb * nb, nb, info,
                    (k + 1) * nb, ((k + 1) & 1) ? WS_DIAG1 : WS_DIAG0, true);
    }
    if (!lookahead)
      potrf_diag<T>(x.s, uplo, tsz(k + 1), blk(k + 1, k + 1), lda, inv + (size_t)((k + 1) & 1) * nb * nb, nb,
                    info, (k + 1) * nb, WS_DIAG0, true);
    keep(k + 1, inv + (size_t)((k + 1) & 1) * nb * nb);
  }
  if (hooks && hooks->after_panel) hooks->after_panel(nt - 1, inv + (size_t)((nt - 1) & 1) * nb * nb);
  if (split) split_cache_clear();
  if (lookahead) gemm_set_priority_stream(nullptr);
  if (own_events) {
    XK_HIP(hipEventDestroy(ev_diag));
    XK_HIP(hipEventDestroy(ev_look));
  }
}

// Diagonal block order of the TRSM sweep for a triangle of order na.
static inline int trsm_block(int nb, int na) { return std::max(64, std::min(nb, na)); }

// dev_trtri over `cnt` triangular blocks of order n at once: block b is
// A + b sA (leading dimension lda), its inverse Tinv + b sInv (ldinv).  Every
// level of the recursion is one batched launch per step for all blocks.
template <typename T>
static void trtri_batched(hipStream_t s, int uplo, int diag, int n, const T* A, long lda, long sA, T* Tinv,
                          long ldinv, long sInv, int cnt, int wslot)
{
  if (n <= 64) {
    trtri_leaf<T>(s, uplo, diag, n, A, lda, sA, Tinv, ldinv, sInv, cnt);
    return;
  }
  const int n1 = half64(n), n2 = n - n1;
  trtri_batched<T>(s, uplo, diag, n1, A, lda, sA, Tinv, ldinv, sInv, cnt, wslot);
  trtri_batched<T>(s, uplo, diag, n2, A + n1 + (long)n1 * lda, lda, sA, Tinv + n1 + (long)n1 * ldinv, ldinv, sInv,
                   cnt, wslot);
  const long sW = (long)n1 * n2;
  T* W = (T*)scratch(wslot, sizeof(T) * (size_t)sW * cnt);
  const T* I11 = Tinv;
  const T* I22 = Tinv + n1 + (long)n1 * ldinv;
  if (uplo == Lower) {
    // I21 = -I22 A21 I11, strictly upper block zeroed
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n1, T(1), A + n1, lda, I11, ldinv, T(0), W, n2, TRI_FULL, cnt, sA, sInv, sW);
    gemm<T>(s, NoTrans, NoTrans, n2, n1, n2, T(-1), I22, ldinv, W, n2, T(0), Tinv + n1, ldinv, TRI_FULL, cnt, sInv,
            sW, sInv);
    T* Z = Tinv + (long)n1 * ldinv;
    gemm<T>(s, NoTrans, NoTrans, n1, n2, 0, T(0), Z, ldinv, Z, ldinv, T(0), Z, ldinv, TRI_FULL, cnt, 0, 0, sInv);
  } else {
    // I12 = -I11 A12 I22, strictly lower block zeroed
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n2, T(1), A + (long)n1 * lda, lda, I22, ldinv, T(0), W, n1, TRI_FULL, cnt,
            sA, sInv, sW);
    gemm<T>(s, NoTrans, NoTrans, n1, n2, n1, T(-1), I11, ldinv, W, n1, T(0), Tinv + (long)n1 * ldinv, ldinv,
            TRI_FULL, cnt, sInv, sW, sInv);
    T* Z = Tinv + n1;
    gemm<T>(s, NoTrans, NoTrans, n2, n1, 0, T(0), Z, ldinv, Z, ldinv, T(0), Z, ldinv, TRI_FULL, cnt, 0, 0, sInv);
  }
}

// Inverses of all nb x nb diagonal blocks of the triangular na x na matrix A,
// stored one after another (nb x nb each) in inv: the full-order blocks in
// one batched recursion, a shorter last block on its own.
template <typename T>
static void trsm_inverses(const Exec& x, int uplo, int diag, int na, const T* A, long lda, int nb, T* inv)
{
  const int full = na / nb, last = na - full * nb;
  if (full > 0)
    trtri_batched<T>(x.s, uplo, diag, nb, A, lda, (long)nb * (1 + lda), inv, nb, (long)nb * nb, full, WS_DIAG0);
  if (last > 0)
    dev_trtri<T>(x, x.s, uplo, diag, last, A + (long)full * nb * (1 + lda), lda, inv + (size_t)full * nb * nb, nb,
                 WS_DIAG0);
}

// The block sweep of dev_trsm with the diagonal-block inverses given.
template <typename T>
static void trsm_sweep(const Exec& x, int side, int uplo, int transA, int m, int n, T alpha, const T* A, long lda,
                       T* B, long ldb, int nb, const T* inv)
{
  const bool tr = transA != NoTrans;
  const bool lower_eff = (uplo == Lower) != tr;
  const int na = side == Left ? m : n;
  const int nblk = (na + nb - 1) / nb;
  auto bsz = [&](int i) { return i == nblk - 1 ? na - i * nb : nb; };
  // op(A) block (i,j) pointer and the transpose flag to use with it
  auto opblk = [&](int i, int j) -> const T* {
    return tr ? A + (long)j * nb + (long)i * nb * lda : A + (long)i * nb + (long)j * nb * lda;
  };
  const int ta = tr ? Trans : NoTrans;
  if (side == Left) {
    T* W = (T*)scratch(WS_TRSMW, sizeof(T) * (size_t)nb * n);
    for (int s = 0; s < nblk; ++s) {
      const int k = lower_eff ? s : nblk - 1 - s;
      const int kk = bsz(k);
      const T la = s == 0 ? alpha : T(1);
      T* Bk = B + (long)k * nb;
      // the inverse product reads Bk and writes it: through W, or through
      // the split-k partials
      if (!gemm_splitk<T>(x.s, ta, NoTrans, kk, n, kk, la, inv + (size_t)k * nb * nb, nb, Bk, ldb, T(0), Bk, ldb)) {
        copy2d<T>(x.s, kk, n, Bk, ldb, W, kk);
        gemm<T>(x.s, ta, NoTrans, kk, n, kk, la, inv + (size_t)k * nb * nb, nb, W, kk, T(0), Bk, ldb);
      }
      if (s == nblk - 1) break;
      const bool plain = !split_enabled<T>(x);
      if (lower_eff) {  // rows below block k
        const int r0 = (k + 1) * nb, rows = m - r0;
        if (!(plain && gemm_splitk<T>(x.s, ta, NoTrans, rows, n, kk, T(-1), opblk(k + 1, k), lda, Bk, ldb, la,
                                      B + r0, ldb)))
          dev_gemm<T>(x, ta, NoTrans, rows, n, kk, T(-1), opblk(k + 1, k), lda, Bk, ldb, la, B + r0, ldb);
      } else {  // rows above block k
        const int rows = k * nb;
        if (!(plain && gemm_splitk<T>(x.s, ta, NoTrans, rows, n, kk, T(-1), opblk(0, k), lda, Bk, ldb, la, B, ldb)))
          dev_gemm<T>(x, ta, NoTrans, rows, n, kk, T(-1), opblk(0, k), lda, Bk, ldb, la, B, ldb);
      }
    }
  } else {
    T* W = (T*)scratch(WS_TRSMW, sizeof(T) * (size_t)m * nb);
    const bool forward = !lower_eff;  // X T = B with T upper: first column block first
    for (int s = 0; s < nblk; ++s) {
      const int k = forward ? s : nblk - 1 - s;
      const int kk = bsz(k);
      const T la = s == 0 ? alpha : T(1);
      T* Bk = B + (long)k * nb * ldb;
      if (!gemm_splitk<T>(x.s, NoTrans, ta, m, kk, kk, la, Bk, ldb, inv + (size_t)k * nb * nb, nb, T(0), Bk, ldb)) {
        copy2d<T>(x.s, m, kk, Bk, ldb, W, m);
        gemm<T>(x.s, NoTrans, ta, m, kk, kk, la, W, m, inv + (size_t)k * nb * nb, nb, T(0), Bk, ldb);
      }
      if (s == nblk - 1) break;
      const bool plain = !split_enabled<T>(x);
      if (forward) {  // columns right of block k
        const int c0 = (k + 1) * nb, cols = n - c0;
        if (!(plain && gemm_splitk<T>(x.s, NoTrans, ta, m, cols, kk, T(-1), Bk, ldb, opblk(k, k + 1), lda, la,
                                      B + (long)c0 * ldb, ldb)))
          dev_gemm<T>(x, NoTrans, ta, m, cols, kk, T(-1), Bk, ldb, opblk(k, k + 1), lda, la, B + (long)c0 * ldb,
                      ldb);
      } else {  // columns left of block k
        const int cols = k * nb;
        if (!(plain && gemm_splitk<T>(x.s, NoTrans, ta, m, cols, kk, T(-1), Bk, ldb, opblk(k, 0), lda, la, B, ldb)))
          dev_gemm<T>(x, NoTrans, ta, m, cols, kk, T(-1), Bk, ldb, opblk(k, 0), lda, la, B, ldb);
      }
    }
  }
}

template <typename T>
void dev_trsm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, T alpha, const T* A, long lda,
              T* B, long ldb, int nb)
{
  if (m <= 0 || n <= 0) return;
  if (alpha == T(0)) {
    fill_zero<T>(x.s, m, n, B, ldb);
    return;
  }
  const int na = side == Left ? m : n;
  nb = trsm_block(nb, na);
  T* inv = (T*)scratch(WS_TRSM, sizeof(T) * (size_t)((na + nb - 1) / nb) * nb * nb);
  trsm_inverses<T>(x, uplo, diag, na, A, lda, nb, inv);
  trsm_sweep<T>(x, side, uplo, transA, m, n, alpha, A, lda, B, ldb, nb, inv);
}

// ------------------------------------------------ GEMMT, SYR2K, SYMM, TRMM
// The reference's tile algorithms for these (blas/zgemmt.c:253-380,
// zsyr2k.c:235-325, zsymm.c:228-394, ztrmm.c:222-460) are GEMM tile loops
// over the output; here each is one or two launches of the same kernels.
enum { WS_X0 = 6, WS_X1 = 7 };

template <typename T>
void dev_gemmt(const Exec& x, int uplo, int transA, int transB, int n, int k, T alpha, const T* A, long lda,
               const T* B, long ldb, T beta, T* C, long ldc)
{
  if (n <= 0) return;
  if (alpha == T(0)) k = 0;  // A and B are not read
  if (uplo != Lower && uplo != Upper) {
    dev_gemm<T>(x, transA, transB, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  const int tri = uplo == Lower ? TRI_LOWER : TRI_UPPER;
  if (split_enabled<T>(x) && k > 0 && alpha != T(0)) {
    split_gemm<T>(x, transA, transB, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri);
    return;
  }
  gemm<T>(x.s, transA, transB, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri);
}

template <typename T>
void dev_syr2k(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, const T* B, long ldb,
               T beta, T* C, long ldc)
{
  // C <- alpha (op(A) op(B)^T + op(B) op(A)^T) + beta C on the uplo triangle
  const int ta = trans == NoTrans ? NoTrans : Trans;
  const int tb = trans == NoTrans ? Trans : NoTrans;
  dev_gemmt<T>(x, uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (k > 0 && alpha != T(0)) dev_gemmt<T>(x, uplo, ta, tb, n, k, alpha, B, ldb, A, lda, T(1), C, ldc);
}

template <typename T>
void dev_symm(const Exec& x, int side, int uplo, int m, int n, T alpha, const T* A, long lda, const T* B, long ldb,
              T beta, T* C, long ldc)
{
  if (m <= 0 || n <= 0) return;
  const int ka = side == Left ? m : n;
  if (alpha == T(0)) {
    dev_gemm<T>(x, NoTrans, NoTrans, m, n, 0, T(0), B, ldb, B, ldb, beta, C, ldc);
    return;
  }
  // the full symmetric operand, then one GEMM
  const long ldf = (ka + 31) / 32 * 32;
  T* F = (T*)scratch(WS_X0, sizeof(T) * (size_t)ldf * ka);
  symm_expand<T>(x.s, uplo, ka, A, lda, F, ldf, nullptr, 0);
  if (side == Left) dev_gemm<T>(x, NoTrans, NoTrans, m, n, m, alpha, F, ldf, B, ldb, beta, C, ldc);
  else dev_gemm<T>(x, NoTrans, NoTrans, m, n, n, alpha, B, ldb, F, ldf, beta, C, ldc);
}

template <typename T>
void dev_trmm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, T alpha, const T* A, long lda,
              T* B, long ldb)
{
  // B <- alpha op(A) B (Left) or alpha B op(A) (Right), A triangular
  if (m <= 0 || n <= 0) return;
  if (alpha == T(0)) {
    fill_zero<T>(x.s, m, n, B, ldb);
    return;
  }
  const int ka = side == Left ? m : n;
  const long ldf = (ka + 31) / 32 * 32, ldw = (m + 31) / 32 * 32;
  T* F = (T*)scratch(WS_X0, sizeof(T) * (size_t)ldf * ka);
  T* W = (T*)scratch(WS_X1, sizeof(T) * (size_t)ldw * n);
  tri_expand<T>(x.s, uplo, diag == Unit, ka, A, lda, F, ldf);
  copy2d<T>(x.s, m, n, B, ldb, W, ldw);
  const int ta = transA == NoTrans ? NoTrans : Trans;
  // op(A) is lower when (uplo == Lower) == (no transpose)
  const bool op_lower = (uplo == Lower) == (ta == NoTrans);
  if (split_enabled<T>(x)) {
    if (side == Left) dev_gemm<T>(x, ta, NoTrans, m, n, m, alpha, F, ldf, W, ldw, T(0), B, ldb);
    else dev_gemm<T>(x, NoTrans, ta, m, n, n, alpha, W, ldw, F, ldf, T(0), B, ldb);
    return;
  }
  // only the inner range where the triangle is not zero
  if (side == Left)
    gemm<T>(x.s, ta, NoTrans, m, n, m, alpha, F, ldf, W, ldw, T(0), B, ldb, TRI_FULL, 1, 0, 0, 0,
            op_lower ? KRANGE_A_LOWER : KRANGE_A_UPPER);
  else
    gemm<T>(x.s, NoTrans, ta, m, n, n, alpha, W, ldw, F, ldf, T(0), B, ldb, TRI_FULL, 1, 0, 0, 0,
            op_lower ? KRANGE_B_LOWER : KRANGE_B_UPPER);
}

// ---------------------------------------------------------------- complex
// [Pr; Pi] = [Xr -Xi; Xi Xr] [Yr; Yi] for X = op(A), Y = op(B): one real
// GEMM (2m x n x 2k); for a (conjugate) transposed operand the transpose of
// its block matrix is built and read transposed.
enum { WS_ZA = 20, WS_ZB = 21, WS_ZP = 22 };

// uplo/herm: write only a triangle of the square C (HERK: real diagonal).
template <typename T>
static void cgemm_impl(const Exec& x, int transA, int transB, int m, int n, int k, const T* alpha, const T* A, long lda,
                       const T* B, long ldb, const T* beta, T* C, long ldc, int uplo, int herm)
{
  if (m <= 0 || n <= 0) return;
  if (k == 0 || (alpha[0] == T(0) && alpha[1] == T(0))) {
    cplx_combine<T>(x.s, m, n, nullptr, 0, alpha, beta, C, ldc, uplo, herm);
    return;
  }
  auto rup = [](long v) { return (v + 31) / 32 * 32; };
  if ((uplo == Lower || uplo == Upper) && m == n) {
    // Only a triangle of C: with X = op(A), Y = op(B) and Z = Y^T,
    //   Pr = Xr Yr - Xi Yi = [Xr Xi] [Zr -Zi]^T,  Pi = Xi Yr + Xr Yi = [Xr Xi] [Zi Zr]^T,
    // two real triangle-masked products of inner dimension 2k: half the
    // flops of the full block product.  Each real operand is packed from the
    // stored matrix without transposing it (a stored k x n matrix gives the
    // stacked 2k x n transpose, which the product reads transposed).
    const bool ta = transA != NoTrans;  // A stored k x n
    const bool tb = transB == NoTrans;  // B stored k x n, so Z = B^T is stored transposed
    const int ar = ta ? k : n, ac = ta ? n : k;
    const int br = tb ? k : n, bc = tb ? n : k;
    const int zi = transB == ConjTrans ? -1 : 1;  // Zi = zi Bi (stored orientation)
    const long ldu = ta ? rup(2L * k) : rup(n), ldv = tb ? rup(2L * k) : rup(n);
    const size_t usz = (size_t)ldu * (ta ? n : 2 * k), vsz = (size_t)ldv * (tb ? n : 2 * k);
    T* U = (T*)scratch(WS_ZA, sizeof(T) * usz);
    T* V = (T*)scratch(WS_ZB, sizeof(T) * 2 * vsz);
    T* V1 = V;
    T* V2 = V + vsz;
    cplx_pack2<T>(x.s, ta, 0, 1, transA == ConjTrans ? -1 : 1, ar, ac, A, lda, U, ldu);
    // MFMA modes: both products in one launch (batch of two: U shared, V1
    // and V2 one stride apart, Pr and Pi n rows apart), so the second fills
    // the partial last round of the first's tiles.  Split modes: two
    // products; [Zr -Zi] equals U when B is A with the conjugate operation
    // (HERK), and the real part is then a SYRK of U
    const bool one_launch = !split_enabled<T>(x);
    const bool same =
        !one_launch && A == B && lda == ldb && ta == tb && (transA == ConjTrans ? -1 : 1) == -zi;
    if (same) V1 = U;
    else cplx_pack2<T>(x.s, tb, 0, 1, -zi, br, bc, B, ldb, V1, ldv);
    cplx_pack2<T>(x.s, tb, 1, 1, zi, br, bc, B, ldb, V2, ldv);
    const long ldp = rup(2L * n);
    T* P = (T*)scratch(WS_ZP, sizeof(T) * (size_t)ldp * n);
    const int opu = ta ? Trans : NoTrans, opv = tb ? NoTrans : Trans;
    if (one_launch) {
      gemm<T>(x.s, opu, opv, n, n, 2 * k, T(1), U, ldu, V1, ldv, T(0), P, ldp, uplo == Lower ? TRI_LOWER : TRI_UPPER,
              2, 0, (long)vsz, (long)n);
    } else {
      dev_gemmt<T>(x, uplo, opu, opv, n, 2 * k, T(1), U, ldu, V1, same ? ldu : ldv, T(0), P, ldp);
      dev_gemmt<T>(x, uplo, opu, opv, n, 2 * k, T(1), U, ldu, V2, ldv, T(0), P + n, ldp);
    }
    cplx_combine<T>(x.s, m, n, P, ldp, alpha, beta, C, ldc, uplo, herm);
    return;
  }
  // op(A): NoTrans m x k -> [Ar -Ai; Ai Ar]; Trans (k x m stored) -> [Ar Ai;
  // -Ai Ar] read transposed; ConjTrans -> [Ar -Ai; Ai Ar] read transposed
  const bool ta = transA != NoTrans;
  const int ar = ta ? k : m, ac = ta ? m : k;
  const long ldya = rup(2L * ar);
  T* YA = (T*)scratch(WS_ZA, sizeof(T) * (size_t)ldya * 2 * ac);
  cplx_pack_a<T>(x.s, transA == Trans ? -1 : 1, ar, ac, A, lda, YA, ldya);
  // op(B): NoTrans -> [Br; Bi]; Trans (n x k stored) -> [Br, Bi] read
  // transposed; ConjTrans -> [Br, -Bi] read transposed
  const bool tb = transB != NoTrans;
  const int br = tb ? n : k, bc = tb ? k : n;
  const long ldyb = rup(tb ? (long)br : 2L * br);
  T* YB = (T*)scratch(WS_ZB, sizeof(T) * (size_t)ldyb * (tb ? 2 * bc : bc));
  cplx_pack_b<T>(x.s, tb ? 0 : 1, transB == ConjTrans ? -1 : 1, br, bc, B, ldb, YB, ldyb);
  const long ldp = rup(2L * m);
  T* P = (T*)scratch(WS_ZP, sizeof(T) * (size_t)ldp * n);
  dev_gemm<T>(x, ta ? Trans : NoTrans, tb ? Trans : NoTrans, 2 * m, n, 2 * k, T(1), YA, ldya, YB, ldyb, T(0), P, ldp);
  cplx_combine<T>(x.s, m, n, P, ldp, alpha, beta, C, ldc, uplo, herm);
}

template <typename T>
void dev_cgemm(const Exec& x, int transA, int transB, int m, int n, int k, const T* alpha, const T* A, long lda,
               const T* B, long ldb, const T* beta, T* C, long ldc)
{
  cgemm_impl<T>(x, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 0);
}

template <typename T>
void dev_cherk(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, T beta, T* C,
               long ldc)
{
  // C <- alpha op(A) op(A)^H + beta C on the uplo triangle, real diagonal
  // (the real part a SYRK of [Ar Ai], the imaginary part one triangle-masked
  // product)
  const T al[2] = {alpha, T(0)}, be[2] = {beta, T(0)};
  const int ta = trans == NoTrans ? NoTrans : ConjTrans, tb = trans == NoTrans ? ConjTrans : NoTrans;
  cgemm_impl<T>(x, ta, tb, n, n, k, al, A, lda, A, lda, be, C, ldc, uplo, 1);
}

template <typename T>
void dev_chemm(const Exec& x, int herm, int side, int uplo, int m, int n, const T* alpha, const T* A, long lda,
               const T* B, long ldb, const T* beta, T* C, long ldc)
{
  const int herm_mode = herm ? CPLX_HERM : CPLX_SYM;
  // the full Hermitian operand, then the complex GEMM
  const int ka = side == Left ? m : n;
  const long ldf = (ka + 15) / 16 * 16;
  T* F = (T*)scratch(WS_X0, 2 * sizeof(T) * (size_t)ldf * ka);
  if (!(alpha[0] == T(0) && alpha[1] == T(0))) cplx_expand<T>(x.s, herm_mode, uplo, ka, A, lda, F, ldf);
  if (side == Left) cgemm_impl<T>(x, NoTrans, NoTrans, m, n, m, alpha, F, ldf, B, ldb, beta, C, ldc, 0, 0);
  else cgemm_impl<T>(x, NoTrans, NoTrans, m, n, n, alpha, B, ldb, F, ldf, beta, C, ldc, 0, 0);
}

template <typename T>
void dev_cgemmt(const Exec& x, int uplo, int transA, int transB, int n, int k, const T* alpha, const T* A, long lda,
                const T* B, long ldb, const T* beta, T* C, long ldc)
{
  cgemm_impl<T>(x, transA, transB, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                uplo == Lower || uplo == Upper ? uplo : 0, 0);
}

template <typename T>
void dev_csyr2k(const Exec& x, int herm, int uplo, int trans, int n, int k, const T* alpha, const T* A, long lda,
                const T* B, long ldb, const T* beta, T* C, long ldc)
{
  // symmetric: alpha (op(A) op(B)^T + op(B) op(A)^T) + beta C; Hermitian:
  // alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C, real diagonal
  const int t = herm ? ConjTrans : Trans;
  const int ta = trans == NoTrans ? NoTrans : t, tb = trans == NoTrans ? t : NoTrans;
  cgemm_impl<T>(x, ta, tb, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, uplo, herm);
  if (k == 0 || (alpha[0] == T(0) && alpha[1] == T(0))) return;
  const T a2[2] = {alpha[0], herm ? -alpha[1] : alpha[1]}, one[2] = {T(1), T(0)};
  cgemm_impl<T>(x, ta, tb, n, n, k, a2, B, ldb, A, lda, one, C, ldc, uplo, herm);
}

template <typename T>
void dev_ctrmm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, const T* alpha, const T* A,
               long lda, T* B, long ldb)
{
  // B <- alpha op(A) B (Left) or alpha B op(A) (Right), A triangular
  const T zero[2] = {T(0), T(0)};
  if (alpha[0] == T(0) && alpha[1] == T(0)) {
    cplx_combine<T>(x.s, m, n, nullptr, 0, zero, zero, B, ldb);
    return;
  }
  const int ka = side == Left ? m : n;
  const long ldf = (ka + 15) / 16 * 16, ldw = (m + 15) / 16 * 16;
  T* F = (T*)scratch(WS_X0, 2 * sizeof(T) * (size_t)ldf * ka);
  T* W = (T*)scratch(WS_X1, 2 * sizeof(T) * (size_t)ldw * n);
  cplx_expand<T>(x.s, diag == Unit ? CPLX_UNIT : CPLX_TRI, uplo, ka, A, lda, F, ldf);
  copy2d<T>(x.s, 2 * m, n, B, 2 * ldb, W, 2 * ldw);  // interleaved: 2 reals per element
  if (side == Left) cgemm_impl<T>(x, transA, NoTrans, m, n, m, alpha, F, ldf, W, ldw, zero, B, ldb, 0, 0);
  else cgemm_impl<T>(x, NoTrans, transA, m, n, n, alpha, W, ldw, F, ldf, zero, B, ldb, 0, 0);
}

// Complex triangular inverse (Tinv n x n, other triangle zeroed): the
// recursion of dev_trtri with complex products; leaves of 64.
enum { WS_ZW = 23 };

template <typename T>
static void dev_ctrtri(const Exec& x, int uplo, int unit, int n, const T* A, long lda, T* Tinv, long ldinv)
{
  if (n <= 64) {
    cplx_trtri_leaf<T>(x.s, uplo, unit, n, A, lda, 0, Tinv, ldinv, 0, 1);
    return;
  }
  const int n1 = half64(n), n2 = n - n1;
  auto at = [](auto* base, long off) { return base + 2 * off; };  // complex element offsets
  const T* A22 = at(A, n1 + (long)n1 * lda);
  T* I22 = at(Tinv, n1 + (long)n1 * ldinv);
  dev_ctrtri<T>(x, uplo, unit, n1, A, lda, Tinv, ldinv);
  dev_ctrtri<T>(x, uplo, unit, n2, A22, lda, I22, ldinv);
  T* W = (T*)scratch(WS_ZW, 2 * sizeof(T) * (size_t)n1 * n2);
  const T one[2] = {T(1), T(0)}, mone[2] = {T(-1), T(0)}, zero[2] = {T(0), T(0)};
  if (uplo == Lower) {
    // I21 = -I22 A21 I11, upper block zeroed
    cgemm_impl<T>(x, NoTrans, NoTrans, n2, n1, n1, one, at(A, n1), lda, Tinv, ldinv, zero, W, n2, 0, 0);
    cgemm_impl<T>(x, NoTrans, NoTrans, n2, n1, n2, mone, I22, ldinv, W, n2, zero, at(Tinv, n1), ldinv, 0, 0);
    cplx_combine<T>(x.s, n1, n2, nullptr, 0, zero, zero, at(Tinv, (long)n1 * ldinv), ldinv);
  } else {
    // I12 = -I11 A12 I22, lower block zeroed
    cgemm_impl<T>(x, NoTrans, NoTrans, n1, n2, n2, one, at(A, (long)n1 * lda), lda, I22, ldinv, zero, W, n1, 0, 0);
    cgemm_impl<T>(x, NoTrans, NoTrans, n1, n2, n1, mone, Tinv, ldinv, W, n1, zero, at(Tinv, (long)n1 * ldinv), ldinv, 0,
                  0);
    cplx_combine<T>(x.s, n2, n1, nullptr, 0, zero, zero, at(Tinv, n1), ldinv);
  }
}

// dev_ctrtri over `cnt` triangular blocks of order n at once (block b at
// A + b sA, its inverse at Tinv + b sInv, strides in complex elements): the
// leaves of every level are one batched launch for all blocks; the
// products run block by block.
template <typename T>
static void ctrtri_batched(const Exec& x, int uplo, int unit, int n, const T* A, long lda, long sA, T* Tinv,
                           long ldinv, long sInv, int cnt)
{
  if (cnt <= 0) return;
  if (n <= 64) {
    cplx_trtri_leaf<T>(x.s, uplo, unit, n, A, lda, sA, Tinv, ldinv, sInv, cnt);
    return;
  }
  const int n1 = half64(n), n2 = n - n1;
  auto at = [](auto* base, long off) { return base + 2 * off; };
  ctrtri_batched<T>(x, uplo, unit, n1, A, lda, sA, Tinv, ldinv, sInv, cnt);
  ctrtri_batched<T>(x, uplo, unit, n2, at(A, n1 + (long)n1 * lda), lda, sA, at(Tinv, n1 + (long)n1 * ldinv), ldinv,
                    sInv, cnt);
  T* W = (T*)scratch(WS_ZW, 2 * sizeof(T) * (size_t)n1 * n2);
  const T one[2] = {T(1), T(0)}, mone[2] = {T(-1), T(0)}, zero[2] = {T(0), T(0)};
  for (int b = 0; b < cnt; ++b) {
    const T* Ab = at(A, (long)b * sA);
    T* Ib = at(Tinv, (long)b * sInv);
    T* I22 = at(Ib, n1 + (long)n1 * ldinv);
    if (uplo == Lower) {
      cgemm_impl<T>(x, NoTrans, NoTrans, n2, n1, n1, one, at(Ab, n1), lda, Ib, ldinv, zero, W, n2, 0, 0);
      cgemm_impl<T>(x, NoTrans, NoTrans, n2, n1, n2, mone, I22, ldinv, W, n2, zero, at(Ib, n1), ldinv, 0, 0);
      cplx_combine<T>(x.s, n1, n2, nullptr, 0, zero, zero, at(Ib, (long)n1 * ldinv), ldinv);
    } else {
      cgemm_impl<T>(x, NoTrans, NoTrans, n1, n2, n2, one, at(Ab, (long)n1 * lda), lda, I22, ldinv, zero, W, n1, 0, 0);
      cgemm_impl<T>(x, NoTrans, NoTrans, n1, n2, n1, mone, Ib, ldinv, W, n1, zero, at(Ib, (long)n1 * ldinv), ldinv, 0,
                    0);
      cplx_combine<T>(x.s, n2, n1, nullptr, 0, zero, zero, at(Ib, n1), ldinv);
    }
  }
}

// Complex TRSM: the block sweep of dev_trsm (inverse products for the
// diagonal blocks, updates by -1 with alpha folded into the first step) with
// complex products; the inverses of all full diagonal blocks come from one
// batched recursion.
template <typename T>
void dev_ctrsm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, const T* alpha, const T* A,
               long lda, T* B, long ldb, int nb)
{
  if (m <= 0 || n <= 0) return;
  const T zero[2] = {T(0), T(0)}, one[2] = {T(1), T(0)}, mone[2] = {T(-1), T(0)};
  if (alpha[0] == T(0) && alpha[1] == T(0)) {
    cplx_combine<T>(x.s, m, n, nullptr, 0, zero, zero, B, ldb);
    return;
  }
  auto at = [](auto* base, long off) { return base + 2 * off; };
  const bool tr = transA != NoTrans;
  const bool lower_eff = (uplo == Lower) != tr;
  const int na = side == Left ? m : n;
  nb = trsm_block(nb, na);
  const int nblk = (na + nb - 1) / nb;
  auto bsz = [&](int i) { return i == nblk - 1 ? na - i * nb : nb; };
  T* inv = (T*)scratch(WS_TRSM, 2 * sizeof(T) * (size_t)nblk * nb * nb);
  const int nfull = bsz(nblk - 1) == nb ? nblk : nblk - 1;
  ctrtri_batched<T>(x, uplo, diag == Unit, nb, A, lda, (long)nb * (1 + lda), inv, nb, (long)nb * nb, nfull);
  if (nfull < nblk)
    dev_ctrtri<T>(x, uplo, diag == Unit, bsz(nblk - 1), at(A, (long)(nblk - 1) * nb * (1 + lda)), lda,
                  at(inv, (long)(nblk - 1) * nb * nb), nb);
  auto opblk = [&](int i, int j) -> const T* {
    return tr ? at(A, (long)j * nb + (long)i * nb * lda) : at(A, (long)i * nb + (long)j * nb * lda);
  };
  if (side == Left) {
    T* W = (T*)scratch(WS_TRSMW, 2 * sizeof(T) * (size_t)nb * n);
    for (int s = 0; s < nblk; ++s) {
      const int k = lower_eff ? s : nblk - 1 - s;
      const int kk = bsz(k);
      const T* la = s == 0 ? alpha : one;
      T* Bk = at(B, (long)k * nb);
      copy2d<T>(x.s, 2 * kk, n, Bk, 2 * ldb, W, 2L * kk);
      cgemm_impl<T>(x, transA, NoTrans, kk, n, kk, la, at(inv, (long)k * nb * nb), nb, W, kk, zero, Bk, ldb, 0, 0);
      if (s == nblk - 1) break;
      if (lower_eff) {
        const int r0 = (k + 1) * nb;
        cgemm_impl<T>(x, transA, NoTrans, m - r0, n, kk, mone, opblk(k + 1, k), lda, Bk, ldb, la, at(B, r0), ldb, 0, 0);
      } else {
        cgemm_impl<T>(x, transA, NoTrans, k * nb, n, kk, mone, opblk(0, k), lda, Bk, ldb, la, B, ldb, 0, 0);
      }
    }
  } else {
    T* W = (T*)scratch(WS_TRSMW, 2 * sizeof(T) * (size_t)m * nb);
    const bool forward = !lower_eff;
    for (int s = 0; s < nblk; ++s) {
      const int k = forward ? s : nblk - 1 - s;
      const int kk = bsz(k);
      const T* la = s == 0 ? alpha : one;
      T* Bk = at(B, (long)k * nb * ldb);
      copy2d<T>(x.s, 2 * m, kk, Bk, 2 * ldb, W, 2L * m);
      cgemm_impl<T>(x, NoTrans, transA, m, kk, kk, la, W, m, at(inv, (long)k * nb * nb), nb, zero, Bk, ldb, 0, 0);
      if (s == nblk - 1) break;
      if (forward) {
        const int c0 = (k + 1) * nb;
        cgemm_impl<T>(x, NoTrans, transA, m, n - c0, kk, mone, Bk, ldb, opblk(k, k + 1), lda, la,
                      at(B, (long)c0 * ldb), ldb, 0, 0);
      } else {
        cgemm_impl<T>(x, NoTrans, transA, m, k * nb, kk, mone, Bk, ldb, opblk(k, 0), lda, la, B, ldb, 0, 0);
      }
    }
  }
}

// Complex Cholesky (A = L L^H lower, U^H U upper): the right-looking blocked
// algorithm of dev_potrf; each diagonal block is factored and inverted
// through its real form, the panel is a product with the inverse, the
// trailing update a Hermitian rank-nb product on the triangle.
enum { WS_ZR = 24, WS_ZPAN = 25, WS_ZINV = 26, WS_ZRI = 29, WS_ZDW = 30 };

// Cholesky factor and inverse of a complex diagonal block of order n through
// its real form: with rho(z) = [re -im; im re], the 2n x 2n real matrix
// M = [rho(a_ij)] is symmetric positive definite when A is Hermitian positive
// definite, and its Cholesky factor is rho(L) (lower triangular, diagonal
// blocks diag(l_ii, l_ii)), whose inverse is rho(inv(L)).  The real recursion
// (64-order multi-wave leaves, one real product per step) replaces a complex
// one whose products are four launches each; the chain is latency-bound, so
// the real form's twice-larger flop count costs little.  R and Rinv hold M,
// then rho(L) and rho(inv(L)) (their lower triangles; Rinv's upper triangle
// is zeroed once by the caller and only ever receives zeros).  Tinv gets
// inv(L) (Lower) or inv(U) = inv(L)^H (Upper) in its uplo triangle.  The
// leading dimension ldr of R and Rinv stays the same for every block, so
// that Rinv's upper triangle stays zero.
template <typename T>
static void cpotrf_diag(hipStream_t s, int uplo, int n, T* A, long lda, T* Tinv, long ldinv, T* R, T* Rinv, long ldr,
                        int* info, int base)
{
  cplx_realform<T>(s, uplo, n, A, lda, R, ldr);
  potrf_diag<T>(s, Lower, 2 * n, R, ldr, Rinv, ldr, info, 2 * base, WS_ZDW, true, 1);
  cplx_fromreal<T>(s, uplo, n, R, ldr, A, lda, Rinv, Tinv, ldinv);
}

// Right-looking blocked Cholesky (the reference's tile algorithm,
// blas/zpotrf.c:89-187, in complex arithmetic) with the structure of
// dev_potrf: each diagonal block is factored and inverted on the side stream
// (its own scratch buffers) while the compute stream applies the previous panel
// to the trailing matrix; the next block column is updated first so that
// its factorization can start at once.
template <typename T>
void dev_cpotrf(const Exec& x, int uplo, int n, T* A, long lda, int nb, int* info)
{
  if (n <= 0) return;
  auto at = [](auto* b, long off) { return b + 2 * off; };
  nb = std::max(64, std::min(nb, n));
  const int nt = (n + nb - 1) / nb;
  const T one[2] = {T(1), T(0)}, mone[2] = {T(-1), T(0)}, zero[2] = {T(0), T(0)};
  // inverses of two consecutive diagonal blocks (ping-pong for lookahead),
  // zeroed once: nothing writes their blocks across the diagonal
  T* inv = (T*)scratch(WS_ZINV, 2 * sizeof(T) * (size_t)2 * nb * nb);
  const bool lookahead = x.s_diag != nullptr && nt > 1;
  hipStream_t sd = lookahead ? x.s_diag : x.s;
  // real forms of one diagonal block and of its factor's inverse: only the
  // diagonal stream uses them, one block after the other
  T* R = (T*)scratch(WS_ZR, sizeof(T) * (size_t)4 * nb * nb);
  T* Rinv = (T*)scratch(WS_ZRI, sizeof(T) * (size_t)4 * nb * nb);
  hipEvent_t ev_diag = x.ev0, ev_look = x.ev1;
  const bool own_events = !ev_diag || !ev_look;
  if (own_events) {
    XK_HIP(hipEventCreateWithFlags(&ev_diag, hipEventDisableTiming));
    XK_HIP(hipEventCreateWithFlags(&ev_look, hipEventDisableTiming));
  }
  if (lookahead) {
    XK_HIP(hipEventRecord(ev_look, x.s));
    XK_HIP(hipStreamWaitEvent(sd, ev_look, 0));
  }
  fill_zero<T>(sd, 2 * nb, 2 * nb, inv, 2L * nb);
  fill_zero<T>(sd, 2 * nb, 2 * nb, Rinv, 2L * nb);
  auto blk = [&](int i, int j) { return at(A, (long)i * nb + (long)j * nb * lda); };
  auto tsz = [&](int i) { return i == nt - 1 ? n - i * nb : nb; };
  cpotrf_diag<T>(sd, uplo, tsz(0), blk(0, 0), lda, inv, nb, R, Rinv, 2L * nb, info, 0);
  for (int k = 0; k < nt; ++k) {
    const int kk = tsz(k);
    T* ik = at(inv, (long)(k & 1) * nb * nb);
    if (lookahead) {
      XK_HIP(hipEventRecord(ev_diag, sd));
      XK_HIP(hipStreamWaitEvent(x.s, ev_diag, 0));
    }
    if (k == nt - 1) break;
    const int rest = n - (k + 1) * nb, k1 = tsz(k + 1), r2 = rest - k1;
    T* W = (T*)scratch(WS_ZPAN, 2 * sizeof(T) * (size_t)kk * rest);
    if (uplo == Lower) {
      T* P = blk(k + 1, k);  // rest x kk panel
      // the panel's rows in block row k+1 first (all that the lookahead
      // reads), then the next diagonal block (lookahead), the rest of the
      // panel, the block column below the next diagonal block and the
      // trailing triangle
      copy2d<T>(x.s, 2 * k1, kk, P, 2 * lda, W, 2L * rest);
      cgemm_impl<T>(x, NoTrans, ConjTrans, k1, kk, kk, one, W, rest, ik, nb, zero, P, lda, 0, 0);
      cgemm_impl<T>(x, NoTrans, ConjTrans, k1, k1, kk, mone, P, lda, P, lda, one, blk(k + 1, k + 1), lda, Lower, 1);
      if (lookahead) XK_HIP(hipEventRecord(ev_look, x.s));
      if (r2 > 0) {
        copy2d<T>(x.s, 2 * r2, kk, at(P, k1), 2 * lda, at(W, k1), 2L * rest);
        cgemm_impl<T>(x, NoTrans, ConjTrans, r2, kk, kk, one, at(W, k1), rest, ik, nb, zero, at(P, k1), lda, 0, 0);
        const T* P2 = at(P, k1);
        cgemm_impl<T>(x, NoTrans, ConjTrans, r2, k1, kk, mone, P2, lda, P, lda, one, blk(k + 2, k + 1), lda, 0, 0);
        cgemm_impl<T>(x, NoTrans, ConjTrans, r2, r2, kk, mone, P2, lda, P2, lda, one, blk(k + 2, k + 2), lda, Lower,
                      1);
      }
    } else {
      T* P = blk(k, k + 1);  // kk x rest panel
      copy2d<T>(x.s, 2 * kk, k1, P, 2 * lda, W, 2L * kk);
      cgemm_impl<T>(x, ConjTrans, NoTrans, kk, k1, kk, one, ik, nb, W, kk, zero, P, lda, 0, 0);
      cgemm_impl<T>(x, ConjTrans, NoTrans, k1, k1, kk, mone, P, lda, P, lda, one, blk(k + 1, k + 1), lda, Upper, 1);
      if (lookahead) XK_HIP(hipEventRecord(ev_look, x.s));
      if (r2 > 0) {
        copy2d<T>(x.s, 2 * kk, r2, at(P, (long)k1 * lda), 2 * lda, at(W, (long)k1 * kk), 2L * kk);
        cgemm_impl<T>(x, ConjTrans, NoTrans, kk, r2, kk, one, ik, nb, at(W, (long)k1 * kk), kk, zero,
                      at(P, (long)k1 * lda), lda, 0, 0);
        const T* P2 = at(P, (long)k1 * lda);
        cgemm_impl<T>(x, ConjTrans, NoTrans, k1, r2, kk, mone, P, lda, P2, lda, one, blk(k + 1, k + 2), lda, 0, 0);
        cgemm_impl<T>(x, ConjTrans, NoTrans, r2, r2, kk, mone, P2, lda, P2, lda, one, blk(k + 2, k + 2), lda, Upper,
                      1);
      }
    }
    T* inext = at(inv, (long)((k + 1) & 1) * nb * nb);
    if (lookahead) XK_HIP(hipStreamWaitEvent(sd, ev_look, 0));
    cpotrf_diag<T>(sd, uplo, k1, blk(k + 1, k + 1), lda, inext, nb, R, Rinv, 2L * nb, info, (k + 1) * nb);
  }
  if (own_events) {
    XK_HIP(hipEventDestroy(ev_diag));
    XK_HIP(hipEventDestroy(ev_look));
  }
}

// ---------------------------------------------------------------- SPD solve
// LAPACK dsposv's mixed-precision scheme (factor in low precision, residuals
// and updates in FP64, fall back to the accurate factor after itermax
// corrections or a failed low-precision factor), with the low-precision
// factor a reduced-moduli split-BF16 Cholesky whose diagonal blocks stay FP64.
// A's full symmetric image is kept for the residual GEMM (B - A X); the
// solves reuse one set of diagonal-block inverses of the factor.
//
// Stopping rule, per column j of x_i = x_{i-1} + d_i (d_i = F^-1 r_{i-1}):
//  * dsposv's backward test on the residual r_{i-1} of the solution the
//    correction was computed from: ||r||_inf <= sqrt(n) ||x||_inf ||A||_inf eps;
//  * and a forward test, because dsposv's sqrt(n) allows a relative error of
//    ~ sqrt(n) eps kappa (4e-14 at n = 32768): the correction is already
//    below IR_TAU ||x||_inf, or, from the second correction on, the observed
//    contraction rho = ||d_i|| / ||d_{i-1}|| predicts the error left in x_i,
//    rho / (1 - rho) ||d_i||, below IR_TAU ||x||_inf, or rho > 1/2 (no further
//    progress: the FP64 residual's own rounding is reached).
// At the correction limit only the backward test is applied (dsposv's).
static constexpr double IR_TAU = 3.552713678800501e-15;  // 2^-48

static bool ir_column_done(bool res_ok, int corr, double dn, double dprev, double xn)
{
  if (!res_ok) return false;
  if (dn <= IR_TAU * xn) return true;
  if (corr < 2 || !(dprev > 0.0)) return false;
  const double rho = dn / dprev;
  if (rho > 0.5) return true;
  return rho / (1.0 - rho) * dn <= IR_TAU * xn;
}

enum { WS_IR_A = 16, WS_IR_F = 17, WS_IR_R = 18, WS_IR_V = 19, WS_IR_P = 28, WS_IR_F32 = 32, WS_IR_S32 = 33 };

int dev_posv_ir(const Exec& x, const Exec& xl, const Exec& xf, int uplo, int n, int nrhs, const double* A, long lda,
                const double* B, long ldb, double* X, long ldx, int nb, int itermax, int* iter)
{
  *iter = 0;
  if (n <= 0 || nrhs <= 0) return 0;
  const long ldf = (n + 31) / 32 * 32;
  // diagonal block order of the solves: few right-hand sides make each sweep
  // step latency-bound, so fewer, larger steps pay (XKBLAS_IR_TRSM_NB)
  static const int ir_nb = getenv("XKBLAS_IR_TRSM_NB") ? atoi(getenv("XKBLAS_IR_TRSM_NB")) : 1024;
  const int tnb = trsm_block(ir_nb > 0 ? ir_nb / 64 * 64 : 1024, n);
  const int nblk = (n + tnb - 1) / tnb;
  double* Af = (double*)scratch(WS_IR_A, sizeof(double) * (size_t)ldf * n);
  double* F = (double*)scratch(WS_IR_F, sizeof(double) * (size_t)ldf * n);
  // R (n x nrhs) then the diagonal-block inverses of the factor
  double* R = (double*)scratch(WS_IR_R, sizeof(double) * ((size_t)ldf * nrhs + (size_t)nblk * tnb * tnb));
  double* inv = R + (size_t)ldf * nrhs;
  // n column sums of |A|, then per column: residual, solution and correction
  // norms; then the info word; then nrhs column maxima (FP32 solves)
  double* V = (double*)scratch(WS_IR_V, sizeof(double) * ((size_t)n + 4 * nrhs + 1));
  int* dinfo = (int*)(V + n + 3 * nrhs);
  std::vector<double> hv((size_t)n + 3 * nrhs), dprev(nrhs, 0.0);
  int hinfo = 0;

  const int tA = uplo == Lower ? NoTrans : Trans;  // op of the first solve's factor
  auto solve = [&](double* Y, long ldy) {          // Y <- F^-1 Y (F = L L^T or U^T U)
    trsm_sweep<double>(x, Left, uplo, tA, n, nrhs, 1.0, F, ldf, Y, ldy, tnb, inv);
    trsm_sweep<double>(x, Left, uplo, tA == NoTrans ? Trans : NoTrans, n, nrhs, 1.0, F, ldf, Y, ldy, tnb, inv);
  };
  // The solves with the reduced factor need only its own accuracy (about
  // 2^-25 with the default moduli): like LAPACK dsposv's single-precision
  // solves they run in FP32 (XKBLAS_IR_SOLVE32=0: FP64), on an FP32 image of
  // the factor formed block column by block column inside the factorization,
  // with each right-hand side scaled by a power of two (no underflow of small
  // residuals).  Only while |A| keeps the factor and the inverses of its
  // diagonal blocks inside FP32's range, and when the factor has more than
  // one diagonal block (one block is factored entirely in FP64: FP32 solves
  // would then cost a correction); the residuals stay accurate FP64.
  static const bool solve32_env = !(getenv("XKBLAS_IR_SOLVE32") && atoi(getenv("XKBLAS_IR_SOLVE32")) == 0);
  const bool want32 = solve32_env && std::max(64, std::min(nb, n)) < n;
  float* F32 = want32 ? (float*)scratch(WS_IR_F32, sizeof(float) * (size_t)ldf * n) : nullptr;
  float* inv32 = want32 ? (float*)scratch(WS_IR_S32, sizeof(float) * ((size_t)nblk * tnb * tnb +
                                                                      (size_t)ldf * nrhs)) : nullptr;
  float* Y32 = want32 ? inv32 + (size_t)nblk * tnb * tnb : nullptr;
  double* cmax = V + n + 3 * nrhs + 1;  // column maxima of the right-hand side being solved
  // Y <- F^-1 Y in FP32 (back: the backward sweep only)
  auto solve32 = [&](double* Y, long ldy, bool back) {
    col_reduce(x.s, 0, n, nrhs, Y, ldy, cmax);
    to_f32(x.s, n, nrhs, Y, ldy, cmax, Y32, ldf);
    if (!back) trsm_sweep<float>(x, Left, uplo, tA, n, nrhs, 1.0f, F32, ldf, Y32, ldf, tnb, inv32);
    trsm_sweep<float>(x, Left, uplo, tA == NoTrans ? Trans : NoTrans, n, nrhs, 1.0f, F32, ldf, Y32, ldf, tnb, inv32);
    from_f32(x.s, n, nrhs, Y32, ldf, cmax, Y, ldy, false);
  };
  // A's full image Af and the factor's copy F: the first block column before
  // the factorization, the rest (and the row sums of |A|, needed only by the
  // stopping test) on the compute stream beside the first diagonal block's
  // factorization, which would otherwise leave it idle
  const int c1 = std::min(n, (std::max(64, std::min(nb, n)) + 63) / 64 * 64);
  symm_expand(x.s, uplo, n, A, lda, Af, ldf, F, ldf, 0, c1);
  auto expand_rest = [&]() {
    symm_expand(x.s, uplo, n, A, lda, Af, ldf, F, ldf, c1, n);
    col_reduce(x.s, 1, n, n, Af, ldf, V);  // row sums of |A| = column sums (symmetric)
  };
  // X <- B: the first solve runs in place, its forward sweep inside the
  // factorization
  copy2d<double>(x.s, n, nrhs, B, ldb, X, ldx);
  // the factorization hands over its diagonal blocks' inverses when the
  // solves' block order divides its own (else they are formed afterwards)
  const int nbf = std::max(64, std::min(nb, n));
  const bool kept = nbf >= n || nbf % tnb == 0;
  // the first solve's forward sweep (L y = b, or U^T y = b), one block of the
  // factorization's order at a time, enqueued on the compute stream as soon
  // as the factor's block column k is final: it runs inside the
  // factorization, whose last steps leave the compute stream waiting for the
  // diagonal chain
  auto forward_step = [&](int k, const double* ik) {
    const int r0 = k * nbf, kk = std::min(nbf, n - r0), rest = n - r0 - kk;
    double* Xk = X + r0;
    // the FP32 image of the factor's block column (Lower) or row (Upper) k
    if (F32) {
      if (uplo == Lower) to_f32(x.s, n - r0, kk, F + r0 + (long)r0 * ldf, ldf, nullptr, F32 + r0 + (long)r0 * ldf, ldf);
      else to_f32(x.s, kk, n - r0, F + r0 + (long)r0 * ldf, ldf, nullptr, F32 + r0 + (long)r0 * ldf, ldf);
    }
    if (!gemm_splitk<double>(x.s, tA, NoTrans, kk, nrhs, kk, 1.0, ik, nbf, Xk, ldx, 0.0, Xk, ldx)) {
      double* W = (double*)scratch(WS_TRSMW, sizeof(double) * (size_t)kk * nrhs);
      copy2d<double>(x.s, kk, nrhs, Xk, ldx, W, kk);
      gemm<double>(x.s, tA, NoTrans, kk, nrhs, kk, 1.0, ik, nbf, W, kk, 0.0, Xk, ldx);
    }
    if (rest > 0) {
      // block column k of L below its diagonal block, or block row k of U
      // right of it (transposed)
      const double* Pk = uplo == Lower ? F + (r0 + kk) + (long)r0 * ldf : F + r0 + (long)(r0 + kk) * ldf;
      if (!gemm_splitk<double>(x.s, tA, NoTrans, rest, nrhs, kk, -1.0, Pk, ldf, Xk, ldx, 1.0, Xk + kk, ldx))
        dev_gemm<double>(x, tA, NoTrans, rest, nrhs, kk, -1.0, Pk, ldf, Xk, ldx, 1.0, Xk + kk, ldx);
    }
  };
  // F <- Cholesky factor of A's image under e; returns info.  first: the
  // reduced factor, with A's remaining expansion beside its first diagonal
  // block and the first solve's forward sweep woven in
  auto factor = [&](const Exec& e, bool first) {
    fill_int(x.s, dinfo, 1, INT_MAX);
    PotrfHooks<double> hk;
    hk.keep_inv = kept ? inv : nullptr;
    hk.keep_nb = tnb;
    if (first) {
      hk.beside_first = expand_rest;
      hk.after_panel = forward_step;
    }
    dev_potrf<double>(e, uplo, n, F, ldf, nb, dinfo, &hk);
    XK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, x.s));
    XK_HIP(hipStreamSynchronize(x.s));
    return hinfo == INT_MAX ? 0 : hinfo;
  };
  double* rn = V + n;            // ||r_j||_inf
  double* xn = V + n + nrhs;     // ||x_j||_inf
  double* dn = V + n + 2 * nrhs; // ||d_j||_inf
  // R <- B - A X with error-free sums of short product runs (the residual's
  // own rounding is what refinement converges to), then the norms of R and X
  void* Pw = scratch(WS_IR_P, resid_scratch_bytes(n, nrhs));
  auto residual = [&]() {
    resid_accurate(x.s, n, nrhs, Af, ldf, X, ldx, B, ldb, R, ldf, Pw);
    col_reduce(x.s, 0, n, nrhs, R, ldf, rn);
    col_reduce(x.s, 0, n, nrhs, X, ldx, xn);
  };
  auto to_host = [&](int count) {  // per-column norms from rn on
    XK_HIP(hipMemcpyAsync(hv.data() + n, rn, sizeof(double) * count * nrhs, hipMemcpyDeviceToHost, x.s));
    XK_HIP(hipStreamSynchronize(x.s));
  };

  double cte = 0.0, anrm = 0.0;
  auto norm_a = [&]() {  // after the factorization's synchronisation
    XK_HIP(hipMemcpyAsync(hv.data(), V, sizeof(double) * n, hipMemcpyDeviceToHost, x.s));
    XK_HIP(hipStreamSynchronize(x.s));
    anrm = 0.0;
    for (int j = 0; j < n; ++j) anrm = (hv[j] != hv[j] || hv[j] > anrm) ? hv[j] : anrm;
    cte = anrm * 1.1102230246251565e-16 * std::sqrt((double)n);  // eps = 2^-53
  };
  // dsposv's test (continue while rnrm > xnrm cte; NaN never passes)
  auto res_ok = [&](int j) { return hv[n + j] <= hv[n + nrhs + j] * cte; };

  int fallback = 0;
  const int finfo = factor(xl, true);
  norm_a();
  if (finfo != 0) {
    fallback = -3;
  } else {
    if (!kept) trsm_inverses<double>(x, uplo, NonUnit, n, F, ldf, tnb, inv);
    const bool s32 = F32 && anrm > 0x1p-100 && anrm < 0x1p+100;
    if (s32) to_f32(x.s, tnb, nblk * tnb, inv, tnb, nullptr, inv32, tnb);
    // the first solve's backward sweep (its forward sweep ran inside the
    // factorization)
    if (s32) solve32(X, ldx, true);
    else trsm_sweep<double>(x, Left, uplo, tA == NoTrans ? Trans : NoTrans, n, nrhs, 1.0, F, ldf, X, ldx, tnb, inv);
    for (int it = 0;; ++it) {
      residual();
      if (it == itermax) {
        to_host(2);
        bool done = true;
        for (int j = 0; j < nrhs && done; ++j) done = res_ok(j);
        if (done) {
          *iter = it;
          return 0;
        }
        break;
      }
      if (s32) solve32(R, ldf, false);  // the correction d
      else solve(R, ldf);
      col_reduce(x.s, 0, n, nrhs, R, ldf, dn);
      add2d(x.s, n, nrhs, R, ldf, X, ldx);
      to_host(3);
      bool done = true;
      for (int j = 0; j < nrhs; ++j) {
        const double d = hv[n + 2 * nrhs + j];
        done = done && ir_column_done(res_ok(j), it + 1, d, dprev[j], hv[n + nrhs + j]);
        dprev[j] = d;
      }
      if (done) {
        *iter = it + 1;
        return 0;
      }
    }
    fallback = -(itermax + 1);
  }
  // the accurate factor, as LAPACK dsposv falls back to dpotrf
  *iter = fallback;
  symm_expand<double>(x.s, uplo, n, A, lda, F, ldf, nullptr, 0);
  const int info = factor(xf, false);
  if (info != 0) return info;
  if (!kept) trsm_inverses<double>(x, uplo, NonUnit, n, F, ldf, tnb, inv);
  copy2d<double>(x.s, n, nrhs, B, ldb, X, ldx);
  solve(X, ldx);
  XK_HIP(hipStreamSynchronize(x.s));
  return 0;
}

#define XK_INST(T)                                                                                             \
  template void dev_gemm<T>(const Exec&, int, int, int, int, int, T, const T*, long, const T*, long, T, T*, long); \
  template void dev_syrk<T>(const Exec&, int, int, int, int, T, const T*, long, T, T*, long);                     \
  template void dev_trsm<T>(const Exec&, int, int, int, int, int, int, T, const T*, long, T*, long, int);        \
  template void dev_potrf<T>(const Exec&, int, int, T*, long, int, int*, const PotrfHooks<T>*);                  \
  template void dev_trtri<T>(const Exec&, hipStream_t, int, int, int, const T*, long, T*, long, int);        \
  template void dev_gemmt<T>(const Exec&, int, int, int, int, int, T, const T*, long, const T*, long, T, T*,     \
                             long);                                                                             \
  template void dev_syr2k<T>(const Exec&, int, int, int, int, T, const T*, long, const T*, long, T, T*, long);  \
  template void dev_symm<T>(const Exec&, int, int, int, int, T, const T*, long, const T*, long, T, T*, long);   \
  template void dev_trmm<T>(const Exec&, int, int, int, int, int, int, T, const T*, long, T*, long);           \
  template void dev_cgemm<T>(const Exec&, int, int, int, int, int, const T*, const T*, long, const T*, long,     \
                             const T*, T*, long);                                                               \
  template void dev_cherk<T>(const Exec&, int, int, int, int, T, const T*, long, T, T*, long);                   \
  template void dev_chemm<T>(const Exec&, int, int, int, int, int, const T*, const T*, long, const T*, long,     \
                             const T*, T*, long);                                                               \
  template void dev_cgemmt<T>(const Exec&, int, int, int, int, int, const T*, const T*, long, const T*, long,    \
                              const T*, T*, long);                                                              \
  template void dev_csyr2k<T>(const Exec&, int, int, int, int, int, const T*, const T*, long, const T*, long,    \
                              const T*, T*, long);                                                              \
  template void dev_ctrmm<T>(const Exec&, int, int, int, int, int, int, const T*, const T*, long, T*, long);     \
  template void dev_ctrsm<T>(const Exec&, int, int, int, int, int, int, const T*, const T*, long, T*, long, int); \
  template void dev_cpotrf<T>(const Exec&, int, int, T*, long, int, int*);
XK_INST(double)
XK_INST(float)
#undef XK_INST

}  // namespace xk
