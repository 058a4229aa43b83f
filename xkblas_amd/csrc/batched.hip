// batched.hip — uneven batches of small GEMMs and lower-triangular TRSMs,
// FP64 and FP32 mixed, in ONE persistent launch (gfx950).
//
// A plan turns the problems into work items — one per 64 x 64 output tile of
// each GEMM, one per 64-column panel of each TRSM — sorted by estimated cost,
// longest first.  The persistent kernel starts as many workgroups as fit on
// the chip at once; each workgroup takes the next item from a device counter
// until the list is exhausted (every workgroup reaches the exit), so large and
// small problems balance without one launch per problem.  A run is two
// launches on one stream: an int fill of the counter (no memset node, so the
// pair can be captured in a hipGraph) and the persistent kernel.
//
// GEMM tile: 4 waves, 2 x 2 of 32 x 32 quarters as 16 x 16 MFMA blocks
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), 16-deep k slices through
// LDS, edges zero-filled.
// TRSM panel (Left, Lower, NoTrans, Non-unit; B m x ncols with ncols <= 64),
// left-looking over 64-row blocks: one MFMA tile forms alpha B_kb minus the
// solved blocks' contribution straight into LDS, four 16-row sub-blocks are
// solved with the solutions in registers (1 / L_ii precomputed), and the
// block is written back for the next one, all inside the workgroup.
// Shared memory is 40 KB per workgroup, so four fit on a CU.
#include "common.h"
#include "kernels.h"
#include "runtime.h"
#include "../../include/xkblas.h"
#include <algorithm>
#include <cstdio>
#include <vector>

namespace xk {

typedef double bd4 __attribute__((ext_vector_type(4)));
typedef float bf4 __attribute__((ext_vector_type(4)));

struct BatchProblem {
  int kind;   // 0 GEMM, 1 TRSM
  int dtype;  // 0 FP64, 1 FP32
  int ta, tb;
  int m, n, k;
  double alpha, beta;
  const void* A;
  long lda;
  const void* B;
  long ldb;
  void* C;
  long ldc;
};

struct BatchItem {
  int prob;
  int tm, tn;  // GEMM tile indices; TRSM: tn = panel index
  int tile;    // GEMM tile order (64 or 32)
};

__device__ __forceinline__ bd4 bmfma(double a, double b, bd4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ bf4 bmfma(float a, float b, bf4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <typename T> struct BAcc;
template <> struct BAcc<double> { typedef bd4 type; };
template <> struct BAcc<float> { typedef bf4 type; };

constexpr int TB_BM = 64, TB_BK = 16, TB_LD = TB_BM + 2;

// Shared memory of one workgroup, in doubles: 40 KB, so that four
// workgroups fit in the 160 KB of a CU.  Overlaid uses:
//   GEMM staging  2 buffers x (A, B) x 16 x 66            4224
//   TRSM          X block 64 x 64 (swizzled) + L strip 64 x 16   5120
//   item index    one int at the start, between items
constexpr int BATCH_SMEM = 5120;

// Thread index hidden from loop-invariant code motion: recomputing the
// swizzled LDS offsets of an item is cheaper than keeping them live (and
// spilled) across the whole persistent loop.
__device__ __forceinline__ int item_tid()
{
  int t = threadIdx.x;
  __asm__ volatile("" : "+v"(t));
  return t;
}

// Main loop of the TBM x TBM tile (TBM = 64 or 32) at (m0, n0) of alpha
// op(A) op(B): 4 waves as 2 x 2 quarters of TBM/2 x TBM/2, each TQ x TQ MFMA
// blocks of 16 x 16 (TQ = TBM / 32); 16-deep k slices double-buffered in
// LDS, the next slice prefetched into registers while the current one is
// multiplied.  Ends with a barrier after the last LDS read, so the caller
// may overwrite `smem` at once.
template <typename T, int TBM>
__device__ __forceinline__ void tile_mainloop(int ta, int tb, int M, int N, int K, const T* __restrict__ A, long lda,
                                              const T* __restrict__ B, long ldb, int m0, int n0, T* smem,
                                              typename BAcc<T>::type (&acc)[TBM / 32][TBM / 32])
{
  constexpr int TQ = TBM / 32, WT = TBM / 2;
  constexpr int V = TBM * TB_BK / 256;  // elements of each slice per thread
  constexpr int SH = TBM == 64 ? 6 : 5;  // log2(TBM)
  T(*As)[TB_BK][TB_LD] = reinterpret_cast<T(*)[TB_BK][TB_LD]>(smem);
  T(*Bs)[TB_BK][TB_LD] = reinterpret_cast<T(*)[TB_BK][TB_LD]>(smem + 2 * TB_BK * TB_LD);
  typedef typename BAcc<T>::type acc_t;
  const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  T ra[V], rb[V];
  auto load = [&](int k0) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int e = tid + 256 * v;  // 0 .. TBM * 16 - 1
      // op(A) element (i, kk): i = e % TBM, kk = e / TBM (coalesced when !ta)
      int i = e & (TBM - 1), kk = e >> SH;
      if (ta) { kk = e & 15; i = e >> 4; }
      const int gi = m0 + i, gk = k0 + kk;
      T x = T(0);
      if (gi < M && gk < K) x = ta ? A[gk + (long)gi * lda] : A[gi + (long)gk * lda];
      ra[v] = x;
      // op(B) element (kk, j)
      int j = e >> 4, kb = e & 15;
      if (tb) { j = e & (TBM - 1); kb = e >> SH; }
      const int gj = n0 + j, gkb = k0 + kb;
      T y = T(0);
      if (gj < N && gkb < K) y = tb ? B[gj + (long)gkb * ldb] : B[gkb + (long)gj * ldb];
      rb[v] = y;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int e = tid + 256 * v;
      int i = e & (TBM - 1), kk = e >> SH;
      if (ta) { kk = e & 15; i = e >> 4; }
      As[buf][kk][i] = ra[v];
      int j = e >> 4, kb = e & 15;
      if (tb) { j = e & (TBM - 1); kb = e >> SH; }
      Bs[buf][kb][j] = rb[v];
    }
  };
#pragma unroll
  for (int i = 0; i < TQ; ++i)
#pragma unroll
    for (int j = 0; j < TQ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  const int nk = (K + TB_BK - 1) / TB_BK;
  const int lr = lane & 15, lk = lane >> 4;
  if (nk > 0) {
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int kt = 0; kt < nk; ++kt) {
      const bool more = kt + 1 < nk;
      if (more) load((kt + 1) * TB_BK);
#pragma unroll
      for (int ks = 0; ks < TB_BK / 4; ++ks) {
        T af[TQ], bf[TQ];
#pragma unroll
        for (int i = 0; i < TQ; ++i) af[i] = As[buf][ks * 4 + lk][wm * WT + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TQ; ++j) bf[j] = Bs[buf][ks * 4 + lk][wn * WT + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < TQ; ++i)
#pragma unroll
          for (int j = 0; j < TQ; ++j) acc[i][j] = bmfma(bf[j], af[i], acc[i][j]);
      }
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
}

template <typename T>
__device__ __forceinline__ void tile64_mainloop(int ta, int tb, int M, int N, int K, const T* __restrict__ A, long lda,
                                                const T* __restrict__ B, long ldb, int m0, int n0, T* smem,
                                                typename BAcc<T>::type (&acc)[2][2])
{
  tile_mainloop<T, 64>(ta, tb, M, N, K, A, lda, B, ldb, m0, n0, smem, acc);
}

// Tile element of accumulator register r of block (i, j) held by `lane`:
// row (of op(A) / C) and column (of op(B) / C) inside a TBM x TBM tile.
template <typename T, int TBM = 64>
__device__ __forceinline__ int tile_row(int wave, int lane, int i)
{
  return (wave & 1) * (TBM / 2) + i * 16 + (lane & 15);
}
template <typename T, int TBM = 64>
__device__ __forceinline__ int tile_col(int wave, int lane, int j, int r)
{
  return (wave >> 1) * (TBM / 2) + j * 16 + (sizeof(T) == 8 ? (lane >> 4) + 4 * r : (lane >> 4) * 4 + r);
}

// C tile (tm, tn) of the TBM x TBM tiling of C <- alpha op(A) op(B) + beta C.
template <typename T, int TBM>
__device__ void gemm_tile(int ta, int tb, int M, int N, int K, T alpha, const T* __restrict__ A, long lda,
                          const T* __restrict__ B, long ldb, T beta, T* __restrict__ C, long ldc, int tm, int tn,
                          T* smem)
{
  constexpr int TQ = TBM / 32;
  typename BAcc<T>::type acc[TQ][TQ];
  const int m0 = tm * TBM, n0 = tn * TBM;
  tile_mainloop<T, TBM>(ta, tb, M, N, K, A, lda, B, ldb, m0, n0, smem, acc);
  const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < TQ; ++i) {
    const int gm = m0 + tile_row<T, TBM>(wave, lane, i);
#pragma unroll
    for (int j = 0; j < TQ; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gn = n0 + tile_col<T, TBM>(wave, lane, j, r);
        if (gm < M && gn < N) {
          T* c = C + gm + (long)gn * ldc;
          T v = alpha * acc[i][j][r];
          if (beta != T(0)) v += beta * *c;
          *c = v;
        }
      }
    }
  }
}

// One 64-column panel of a lower-triangular solve L X = alpha B, in place,
// left-looking over 64-row blocks kb:
//   X_kb <- alpha B_kb - L[kb, 0:kb] X[0:kb]   one MFMA tile, k = 64 kb, into LDS
//   X_kb <- L_kk^-1 X_kb                       in LDS, 16-row sub-blocks
//   X_kb -> global                             (read back by later blocks)
// The X block lives in LDS as Xs[i * 64 + (c ^ i)] (row i, column c): rows
// and columns are both conflict-free permutations of the banks.  L_kk's lower
// triangle is read into registers before the block's GEMM and handed to LDS
// one 16-column strip at a time: Ld[i * 16 + (q ^ (i & 15))] = L(i, 16 s + q),
// with 1 / L(i, i) in place of the diagonal.  Per strip s, wave 0 (lane =
// column) solves rows 16 s .. 16 s + 15 with the 16 solutions in registers,
// then all four waves subtract their contribution from the rows below.
template <typename T>
__device__ void trsm_panel(const BatchProblem& p, int panel, T* smem)
{
  const int m = p.m;
  const int c0 = panel * 64;
  const int nc = min(64, p.n - c0);
  const T* L = reinterpret_cast<const T*>(p.A);
  T* Bp = reinterpret_cast<T*>(p.C) + (long)c0 * p.ldc;
  const long ldl = p.lda, ldb = p.ldc;
  const T alpha = (T)p.alpha;
  T* Xs = smem;
  T* Ld = smem + 64 * 64;
  const int nb = (m + 63) / 64;
  for (int kb = 0; kb < nb; ++kb) {
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int r0 = kb * 64, h = min(64, m - r0);
    // L_kk(lane, 4 v + wave), lower triangle; padding rows get a unit diagonal
    T lreg[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int q = 4 * v + wave;
      T x = T(0);
      if (lane < h && q < lane) x = L[(r0 + lane) + (long)(r0 + q) * ldl];
      else if (q == lane) x = lane < h ? L[(r0 + lane) + (long)(r0 + q) * ldl] : T(1);
      lreg[v] = x;
    }
    typename BAcc<T>::type acc[2][2];
    tile64_mainloop<T>(0, 0, h, nc, r0, L + r0, ldl, Bp, ldb, 0, 0, smem, acc);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = tile_row<T>(wave, lane, i);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = tile_col<T>(wave, lane, j, r);
          T v = T(0);
          if (row < h && col < nc) v = alpha * Bp[(r0 + row) + (long)col * ldb] - acc[i][j][r];
          Xs[row * 64 + (col ^ row)] = v;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (16 * s >= h) break;
      // strip s of L_kk: this thread holds columns q = 4 v + wave, v = 4 s .. 4 s + 3
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = 4 * u + wave;  // column inside the strip
        T x = lreg[4 * s + u];
        if (16 * s + q == lane) x = T(1) / x;
        Ld[lane * 16 + (q ^ (lane & 15))] = x;
      }
      __syncthreads();
      if (wave == 0) {
        const int c = lane;
        T x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = Xs[(16 * s + r) * 64 + (c ^ (16 * s + r))];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const T* lrow = Ld + (16 * s + r) * 16;
          T v = x[r];
#pragma unroll
          for (int q = 0; q < r; ++q) v -= lrow[q ^ r] * x[q];
          x[r] = v * lrow[0];  // lrow[r ^ r] = 1 / L(r, r)
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) Xs[(16 * s + r) * 64 + (c ^ (16 * s + r))] = x[r];
      }
      __syncthreads();
      if (16 * (s + 1) < h) {
        const int c = lane;
        T xq[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) xq[q] = Xs[(16 * s + q) * 64 + (c ^ (16 * s + q))];
        for (int i = 16 * (s + 1) + wave; i < h; i += 4) {
          const T* lrow = Ld + i * 16;
          T v = Xs[i * 64 + (c ^ i)];
#pragma unroll
          for (int q = 0; q < 16; ++q) v -= lrow[q ^ (i & 15)] * xq[q];
          Xs[i * 64 + (c ^ i)] = v;
        }
      }
      __syncthreads();
    }
    // X_kb back to global memory, coalesced along rows
#pragma unroll 4
    for (int v = 0; v < 16; ++v) {
      const int i = lane, c = wave + 4 * v;
      if (i < h && c < nc) Bp[(r0 + i) + (long)c * ldb] = Xs[i * 64 + (c ^ i)];
    }
    __syncthreads();  // X_kb visible to the workgroup's next block; Xs free
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void batch_kernel(
    const BatchProblem* __restrict__ probs, const BatchItem* __restrict__ items, int nitems, int* counter)
{
  __shared__ double smem_d[BATCH_SMEM];
  int* next = reinterpret_cast<int*>(smem_d);
  for (;;) {
    if (threadIdx.x == 0) *next = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(*next);
    __syncthreads();
    if (it >= nitems) return;
    const BatchItem w = items[it];
    const BatchProblem p = probs[w.prob];
    if (p.kind == 0) {
      const int ta = p.ta != NoTrans, tb = p.tb != NoTrans;
      if (p.dtype == 0) {
        const double* A = reinterpret_cast<const double*>(p.A);
        const double* B = reinterpret_cast<const double*>(p.B);
        double* C = reinterpret_cast<double*>(p.C);
        if (w.tile == 32)
          gemm_tile<double, 32>(ta, tb, p.m, p.n, p.k, p.alpha, A, p.lda, B, p.ldb, p.beta, C, p.ldc, w.tm, w.tn,
                                smem_d);
        else
          gemm_tile<double, 64>(ta, tb, p.m, p.n, p.k, p.alpha, A, p.lda, B, p.ldb, p.beta, C, p.ldc, w.tm, w.tn,
                                smem_d);
      } else {
        const float* A = reinterpret_cast<const float*>(p.A);
        const float* B = reinterpret_cast<const float*>(p.B);
        float* C = reinterpret_cast<float*>(p.C);
        float* sm = reinterpret_cast<float*>(smem_d);
        if (w.tile == 32)
          gemm_tile<float, 32>(ta, tb, p.m, p.n, p.k, (float)p.alpha, A, p.lda, B, p.ldb, (float)p.beta, C, p.ldc,
                               w.tm, w.tn, sm);
        else
          gemm_tile<float, 64>(ta, tb, p.m, p.n, p.k, (float)p.alpha, A, p.lda, B, p.ldb, (float)p.beta, C, p.ldc,
                               w.tm, w.tn, sm);
      }
    } else {
      // a TRSM panel is a chain of dependent blocks and the longest ones set
      // the batch's time: its waves issue ahead of the GEMM items beside them
      __builtin_amdgcn_s_setprio(2);
      if (p.dtype == 0) trsm_panel<double>(p, w.tn, smem_d);
      else trsm_panel<float>(p, w.tn, reinterpret_cast<float*>(smem_d));
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();  // smem free for the next item index
  }
}

// ---------------------------------------------------------------- plan
struct BatchPlan {
  BatchProblem* d_probs = nullptr;
  BatchItem* d_items = nullptr;
  int* d_counter = nullptr;
  int nitems = 0, nprobs = 0;
  int grid = 0;
  double flops = 0.0, bytes = 0.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void enqueue_run(BatchPlan* pl, hipStream_t s)
{
  fill_int(s, pl->d_counter, 1, 0);
  hipLaunchKernelGGL(batch_kernel, dim3(pl->grid), dim3(256), 0, s, pl->d_probs, pl->d_items, pl->nitems,
                     pl->d_counter);
  XK_HIP(hipGetLastError());
}

// Work list of a batch: one item per 64 x 64 GEMM tile and per 64-column
// TRSM panel, sorted by estimated cost, longest first (ties keep problem
// order).  batch_count_items counts the items; batch_build fills hp[count]
// and work[], returns the item count and adds the batch's flops and bytes.
struct BatchCost {
  double c;
  long seq;
  BatchItem w;
};

// GEMM item shape: 1 = 32 x 32 tiles for a problem whose 32 x 32 tiling
// covers at most 3/4 of the area the 64 x 64 one does (a small m or n), 2 =
// 32 x 32 only, 3 = 64 x 64 only, 0 (default) = rule 1 for FP64 problems and
// 64 x 64 for FP32 ones (XKBLAS_BATCH_TILE).  On the uneven batch of 4096
// GEMMs + 1024 TRSMs (MI355X, best of five per run, four runs), FP64: per
// problem 1.54-1.70 ms (mean 1.60) vs 64 x 64 1.56-1.80 (mean 1.70); FP32:
// 64 x 64 1.06-1.10 vs per problem 1.13-1.14 ms.
static int g_batch_tile = getenv("XKBLAS_BATCH_TILE") ? atoi(getenv("XKBLAS_BATCH_TILE")) : 0;

static int gemm_item_tile(int m, int n, int dtype)
{
  const int mode = g_batch_tile == 0 ? (dtype == 0 ? 1 : 3) : g_batch_tile;
  if (mode == 3) return 64;
  if (mode == 2) return 32;
  const long a64 = (long)((m + 63) / 64) * ((n + 63) / 64) * 4096;
  const long a32 = (long)((m + 31) / 32) * ((n + 31) / 32) * 1024;
  return 4 * a32 <= 3 * a64 ? 32 : 64;
}

static long batch_count_items(const xkblas_batch_problem_t* probs, int count)
{
  long n = 0;
  for (int i = 0; i < count; ++i) {
    const xkblas_batch_problem_t& q = probs[i];
    if (q.m <= 0 || q.n <= 0) continue;
    if (q.kind == XKBLAS_BATCH_GEMM) {
      const int t = gemm_item_tile(q.m, q.n, q.dtype);
      n += (long)((q.m + t - 1) / t) * ((q.n + t - 1) / t);
    } else {
      n += (q.n + 63) / 64;
    }
  }
  return n;
}

static long batch_build(const xkblas_batch_problem_t* probs, int count, BatchProblem* hp, BatchCost* work,
                        double* flops, double* bytes)
{
  long nw = 0;
  for (int i = 0; i < count; ++i) {
    const xkblas_batch_problem_t& q = probs[i];
    BatchProblem& p = hp[i];
    p.kind = q.kind; p.dtype = q.dtype; p.ta = q.transA; p.tb = q.transB;
    p.m = q.m; p.n = q.n; p.k = q.k; p.alpha = q.alpha; p.beta = q.beta;
    p.A = q.A; p.lda = q.lda; p.B = q.B; p.ldb = q.ldb; p.C = q.C; p.ldc = q.ldc;
    const double elt = q.dtype == 0 ? 8.0 : 4.0;
    XK_CHECK(q.kind == XKBLAS_BATCH_GEMM || q.kind == XKBLAS_BATCH_TRSM, "batch: unknown problem kind");
    XK_CHECK(q.dtype == 0 || q.dtype == 1, "batch: dtype must be 0 (FP64) or 1 (FP32)");
    if (q.m <= 0 || q.n <= 0) continue;
    if (q.kind == XKBLAS_BATCH_GEMM) {
      const int t = gemm_item_tile(q.m, q.n, q.dtype);
      const int tmx = (q.m + t - 1) / t, tnx = (q.n + t - 1) / t;
      // a tile's k-loop work plus a fixed overhead (a 32 x 32 tile runs at a
      // quarter of the MFMA work per slice, hence the same time per k as half
      // of a 64 x 64 one)
      const double c = t == 64 ? 2.0 * 64 * 64 * std::max(q.k, 1) + 64 * 64 * 8.0
                               : 2.0 * 64 * 32 * std::max(q.k, 1) + 64 * 64 * 8.0;
      for (int a = 0; a < tmx; ++a)
        for (int b = 0; b < tnx; ++b) {
          work[nw] = BatchCost{c, nw, BatchItem{i, a, b, t}};
          ++nw;
        }
      *flops += 2.0 * q.m * q.n * (double)q.k;
      *bytes += elt * ((double)q.m * q.k + (double)q.k * q.n + 2.0 * q.m * q.n);
    } else {
      const int np = (q.n + 63) / 64;
      // left-looking blocks: one tile of k = 64 kb each plus the block's
      // solve, priced like a tile of k = 128
      const double nbk = (q.m + 63) / 64;
      const double c = 2.0 * 64 * 64 * 64 * nbk * (nbk - 1) / 2 + nbk * 2.0 * 64 * 64 * 128;
      for (int b = 0; b < np; ++b) {
        work[nw] = BatchCost{c, nw, BatchItem{i, 0, b, 64}};
        ++nw;
      }
      *flops += (double)q.m * q.m * q.n;
      *bytes += elt * (0.5 * (double)q.m * q.m + 2.0 * q.m * q.n);
    }
  }
  std::sort(work, work + nw,
            [](const BatchCost& a, const BatchCost& b) { return a.c > b.c || (a.c == b.c && a.seq < b.seq); });
  return nw;
}

// workgroups of the persistent kernel: as many as can be resident at once
static int batch_grid(long nitems)
{
  int dev = 0, cus = 0, per_cu = 0;
  XK_HIP(hipGetDevice(&dev));
  XK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  XK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel, 256, 0));
  return (int)std::max(1L, std::min((long)cus * std::max(per_cu, 1), std::max(nitems, 1L)));
}

// ---- allocation-free runs
// xkblas_batch_run_noalloc serves callers that cannot allocate device
// memory: its counter, problem list and work list live in one fixed device
// array (`scratch`, below; one copy per GPU, held until the process ends),
// whose device address hipGetSymbolAddress gives at initialisation.  The
// host-side lists are built in fixed host arrays, so the call allocates
// nothing on the host either.  Calls share the array: they are ordered on
// the library's compute stream, and calls that overlap on one GPU from
// other streams would race.
constexpr int NOALLOC_MAX_PROBS = 8192;
constexpr long NOALLOC_MAX_ITEMS = 196608;
constexpr size_t NOALLOC_PROBS_OFF = 256;  // the work counter comes first
constexpr size_t NOALLOC_ITEMS_OFF = NOALLOC_PROBS_OFF + sizeof(BatchProblem) * NOALLOC_MAX_PROBS;
constexpr size_t NOALLOC_BYTES = NOALLOC_ITEMS_OFF + sizeof(BatchItem) * NOALLOC_MAX_ITEMS;
constexpr int NOALLOC_MAX_DEVICES = 64;

static char* g_noalloc_dev[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per GPU
static BatchProblem g_noalloc_probs[NOALLOC_MAX_PROBS];
static BatchCost g_noalloc_work[NOALLOC_MAX_ITEMS];
static BatchItem g_noalloc_items[NOALLOC_MAX_ITEMS];
static hipEvent_t g_noalloc_staged = nullptr;  // the last run's copies have read the host lists

}  // namespace xk

// 3.9 MB: counter, NOALLOC_MAX_PROBS problems, NOALLOC_MAX_ITEMS work items.
// In host code this name is only the symbol's host-side handle: kernels get
// the address hipGetSymbolAddress returns.
static __device__ __align__(256) char scratch[xk::NOALLOC_BYTES];

namespace xk {

void batch_noalloc_init()
{
  int dev = 0;
  XK_HIP(hipGetDevice(&dev));
  XK_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device ordinal beyond the allocation-free table");
  if (g_noalloc_dev[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p) fatal("hipGetSymbolAddress(scratch)", __FILE__, __LINE__, hipGetErrorString(e));
  g_noalloc_dev[dev] = static_cast<char*>(p);
  if (!g_noalloc_staged) XK_HIP(hipEventCreateWithFlags(&g_noalloc_staged, hipEventDisableTiming));
}

}  // namespace xk

using namespace xk;

extern "C" {

void* xkblas_batch_create(const xkblas_batch_problem_t* probs, int count)
{
  if (count < 0 || (count > 0 && !probs)) return nullptr;
  xkblas_init();
  BatchPlan* pl = new BatchPlan();
  std::vector<BatchProblem> hp(std::max(count, 1));
  std::vector<BatchCost> work(std::max(batch_count_items(probs, count), 1L));
  const long nw = batch_build(probs, count, hp.data(), work.data(), &pl->flops, &pl->bytes);
  std::vector<BatchItem> items(std::max(nw, 1L));
  for (long i = 0; i < nw; ++i) items[i] = work[i].w;
  pl->nitems = (int)nw;
  pl->nprobs = count;
  XK_HIP(hipMalloc(&pl->d_probs, sizeof(BatchProblem) * std::max(count, 1)));
  XK_HIP(hipMalloc(&pl->d_items, sizeof(BatchItem) * items.size()));
  XK_HIP(hipMalloc(&pl->d_counter, sizeof(int)));
  if (count) XK_HIP(hipMemcpy(pl->d_probs, hp.data(), sizeof(BatchProblem) * count, hipMemcpyHostToDevice));
  if (nw) XK_HIP(hipMemcpy(pl->d_items, items.data(), sizeof(BatchItem) * nw, hipMemcpyHostToDevice));
  pl->grid = batch_grid(nw);
  return pl;
}

int xkblas_batch_run(void* plan)
{
  BatchPlan* pl = static_cast<BatchPlan*>(plan);
  if (!pl) return -1;
  std::lock_guard<std::mutex> lock(api_mutex());
  Runtime& r = rt();
  r.before_compute();
  r.prof_begin("batch", pl->flops);
  if (pl->exec) XK_HIP(hipGraphLaunch(pl->exec, r.s_comp));
  else enqueue_run(pl, r.s_comp);
  r.prof_end();
  return 0;
}

int xkblas_batch_run_noalloc(const xkblas_batch_problem_t* problems, int count)
{
  if (count < 0 || (count > 0 && !problems)) return -1;
  if (count > NOALLOC_MAX_PROBS) {
    fprintf(stderr, "[xkblas_amd] xkblas_batch_run_noalloc: %d problems exceed the maximum of %d\n", count,
            NOALLOC_MAX_PROBS);
    return -2;
  }
  const long nw = batch_count_items(problems, count);
  if (nw > NOALLOC_MAX_ITEMS) {
    fprintf(stderr, "[xkblas_amd] xkblas_batch_run_noalloc: %ld work items exceed the maximum of %ld\n", nw,
            NOALLOC_MAX_ITEMS);
    return -3;
  }
  std::lock_guard<std::mutex> lock(api_mutex());
  Runtime& r = rt();
  if (!r.device_ready) {
    fprintf(stderr, "[xkblas_amd] xkblas_batch_run_noalloc: call xkblas_init first\n");
    return -4;
  }
  int dev = 0;
  XK_HIP(hipGetDevice(&dev));
  XK_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES && g_noalloc_dev[dev], "allocation-free array not set up");
  char* base = g_noalloc_dev[dev];
  // the previous run's copies must have read the host lists before they change
  XK_HIP(hipEventSynchronize(g_noalloc_staged));
  double flops = 0.0, bytes = 0.0;
  batch_build(problems, count, g_noalloc_probs, g_noalloc_work, &flops, &bytes);
  for (long i = 0; i < nw; ++i) g_noalloc_items[i] = g_noalloc_work[i].w;
  BatchProblem* d_probs = reinterpret_cast<BatchProblem*>(base + NOALLOC_PROBS_OFF);
  BatchItem* d_items = reinterpret_cast<BatchItem*>(base + NOALLOC_ITEMS_OFF);
  int* d_counter = reinterpret_cast<int*>(base);
  r.before_compute();
  if (count)
    XK_HIP(hipMemcpyAsync(d_probs, g_noalloc_probs, sizeof(BatchProblem) * count, hipMemcpyHostToDevice, r.s_comp));
  if (nw) XK_HIP(hipMemcpyAsync(d_items, g_noalloc_items, sizeof(BatchItem) * nw, hipMemcpyHostToDevice, r.s_comp));
  XK_HIP(hipEventRecord(g_noalloc_staged, r.s_comp));
  fill_int(r.s_comp, d_counter, 1, 0);
  hipLaunchKernelGGL(batch_kernel, dim3(batch_grid(nw)), dim3(256), 0, r.s_comp, d_probs, d_items, (int)nw,
                     d_counter);
  XK_HIP(hipGetLastError());
  return 0;
}

void xkblas_set_batch_tile(int mode)
{
  std::lock_guard<std::mutex> lock(api_mutex());
  g_batch_tile = mode < 0 || mode > 3 ? 0 : mode;
}

int xkblas_batch_use_graph(void* plan, int on)
{
  BatchPlan* pl = static_cast<BatchPlan*>(plan);
  if (!pl) return -1;
  std::lock_guard<std::mutex> lock(api_mutex());
  Runtime& r = rt();
  if (pl->exec) {
    XK_HIP(hipGraphExecDestroy(pl->exec));
    XK_HIP(hipGraphDestroy(pl->graph));
    pl->exec = nullptr;
    pl->graph = nullptr;
  }
  if (!on) return 0;
  XK_HIP(hipStreamSynchronize(r.s_comp));
  XK_HIP(hipStreamBeginCapture(r.s_comp, hipStreamCaptureModeThreadLocal));
  enqueue_run(pl, r.s_comp);
  XK_HIP(hipStreamEndCapture(r.s_comp, &pl->graph));
  XK_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  return 0;
}

int xkblas_batch_info(void* plan, int* nitems, int* grid, double* flops, double* bytes)
{
  BatchPlan* pl = static_cast<BatchPlan*>(plan);
  if (!pl) return -1;
  if (nitems) *nitems = pl->nitems;
  if (grid) *grid = pl->grid;
  if (flops) *flops = pl->flops;
  if (bytes) *bytes = pl->bytes;
  return 0;
}

void xkblas_batch_destroy(void* plan)
{
  BatchPlan* pl = static_cast<BatchPlan*>(plan);
  if (!pl) return;
  XK_HIP(hipDeviceSynchronize());
  if (pl->exec) {
    (void)hipGraphExecDestroy(pl->exec);
    (void)hipGraphDestroy(pl->graph);
  }
  (void)hipFree(pl->d_probs);
  (void)hipFree(pl->d_items);
  (void)hipFree(pl->d_counter);
  delete pl;
}

void* xkblas_device_malloc(size_t bytes)
{
  xkblas_init();
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
  return p;
}

void xkblas_device_free(void* p)
{
  if (p) {
    (void)hipDeviceSynchronize();
    (void)hipFree(p);
  }
}

int xkblas_memcpy_htod(void* dst, const void* src, size_t bytes)
{
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int xkblas_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

}  // extern "C"
