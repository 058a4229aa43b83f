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
  int pad;
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

// Main loop of the 64 x 64 tile (tm, tn) of alpha op(A) op(B): 4 waves as
// 2 x 2 quarters of 32 x 32, each 2 x 2 MFMA blocks of 16 x 16; 16-deep k
// slices double-buffered in LDS, the next slice prefetched into registers
// while the current one is multiplied.  Ends with a barrier after the last
// LDS read, so the caller may overwrite `smem` at once.
template <typename T>
__device__ __forceinline__ void tile64_mainloop(int ta, int tb, int M, int N, int K, const T* __restrict__ A, long lda,
                                                const T* __restrict__ B, long ldb, int m0, int n0, T* smem,
                                                typename BAcc<T>::type (&acc)[2][2])
{
  T(*As)[TB_BK][TB_LD] = reinterpret_cast<T(*)[TB_BK][TB_LD]>(smem);
  T(*Bs)[TB_BK][TB_LD] = reinterpret_cast<T(*)[TB_BK][TB_LD]>(smem + 2 * TB_BK * TB_LD);
  typedef typename BAcc<T>::type acc_t;
  const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  // each thread moves 4 elements of each 64 x 16 slice of op(A) and op(B)
  T ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = tid + 256 * v;  // 0 .. 1023
      // op(A) element (i, kk): i = e % 64, kk = e / 64 (coalesced when !ta)
      int i = e & 63, kk = e >> 6;
      if (ta) { kk = e & 15; i = e >> 4; }
      const int gi = m0 + i, gk = k0 + kk;
      T x = T(0);
      if (gi < M && gk < K) x = ta ? A[gk + (long)gi * lda] : A[gi + (long)gk * lda];
      ra[v] = x;
      // op(B) element (kk, j)
      int j = e >> 4, kb = e & 15;
      if (tb) { j = e & 63; kb = e >> 6; }
      const int gj = n0 + j, gkb = k0 + kb;
      T y = T(0);
      if (gj < N && gkb < K) y = tb ? B[gj + (long)gkb * ldb] : B[gkb + (long)gj * ldb];
      rb[v] = y;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = tid + 256 * v;
      int i = e & 63, kk = e >> 6;
      if (ta) { kk = e & 15; i = e >> 4; }
      As[buf][kk][i] = ra[v];
      int j = e >> 4, kb = e & 15;
      if (tb) { j = e & 63; kb = e >> 6; }
      Bs[buf][kb][j] = rb[v];
    }
  };
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
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
        T af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = As[buf][ks * 4 + lk][wm * 32 + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < 2; ++j) bf[j] = Bs[buf][ks * 4 + lk][wn * 32 + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = bmfma(bf[j], af[i], acc[i][j]);
      }
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
}

// Tile element of accumulator register r of block (i, j) held by `lane`:
// row (of op(A) / C) and column (of op(B) / C) inside the 64 x 64 tile.
template <typename T>
__device__ __forceinline__ int tile_row(int wave, int lane, int i) { return (wave & 1) * 32 + i * 16 + (lane & 15); }
template <typename T>
__device__ __forceinline__ int tile_col(int wave, int lane, int j, int r)
{
  return (wave >> 1) * 32 + j * 16 + (sizeof(T) == 8 ? (lane >> 4) + 4 * r : (lane >> 4) * 4 + r);
}

// C tile (tm, tn) of the 64 x 64 tiling of C <- alpha op(A) op(B) + beta C.
template <typename T>
__device__ void gemm_tile64(int ta, int tb, int M, int N, int K, T alpha, const T* __restrict__ A, long lda,
                            const T* __restrict__ B, long ldb, T beta, T* __restrict__ C, long ldc, int tm, int tn,
                            T* smem)
{
  typename BAcc<T>::type acc[2][2];
  const int m0 = tm * TB_BM, n0 = tn * TB_BM;
  tile64_mainloop<T>(ta, tb, M, N, K, A, lda, B, ldb, m0, n0, smem, acc);
  const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gm = m0 + tile_row<T>(wave, lane, i);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gn = n0 + tile_col<T>(wave, lane, j, r);
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
      if (p.dtype == 0)
        gemm_tile64<double>(p.ta != NoTrans, p.tb != NoTrans, p.m, p.n, p.k, p.alpha,
                            reinterpret_cast<const double*>(p.A), p.lda, reinterpret_cast<const double*>(p.B), p.ldb,
                            p.beta, reinterpret_cast<double*>(p.C), p.ldc, w.tm, w.tn, smem_d);
      else
        gemm_tile64<float>(p.ta != NoTrans, p.tb != NoTrans, p.m, p.n, p.k, (float)p.alpha,
                           reinterpret_cast<const float*>(p.A), p.lda, reinterpret_cast<const float*>(p.B), p.ldb,
                           (float)p.beta, reinterpret_cast<float*>(p.C), p.ldc, w.tm, w.tn,
                           reinterpret_cast<float*>(smem_d));
    } else {
      if (p.dtype == 0) trsm_panel<double>(p, w.tn, smem_d);
      else trsm_panel<float>(p, w.tn, reinterpret_cast<float*>(smem_d));
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

}  // namespace xk

using namespace xk;

extern "C" {

void* xkblas_batch_create(const xkblas_batch_problem_t* probs, int count)
{
  if (count < 0 || (count > 0 && !probs)) return nullptr;
  xkblas_init();
  BatchPlan* pl = new BatchPlan();
  std::vector<BatchProblem> hp(count);
  struct Cost {
    double c;
    BatchItem w;
  };
  std::vector<Cost> work;
  for (int i = 0; i < count; ++i) {
    const xkblas_batch_problem_t& q = probs[i];
    BatchProblem& p = hp[i];
    p.kind = q.kind; p.dtype = q.dtype; p.ta = q.transA; p.tb = q.transB;
    p.m = q.m; p.n = q.n; p.k = q.k; p.alpha = q.alpha; p.beta = q.beta;
    p.A = q.A; p.lda = q.lda; p.B = q.B; p.ldb = q.ldb; p.C = q.C; p.ldc = q.ldc;
    const double elt = q.dtype == 0 ? 8.0 : 4.0;
    XK_CHECK(q.kind == XKBLAS_BATCH_GEMM || q.kind == XKBLAS_BATCH_TRSM, "batch: unknown problem kind");
    XK_CHECK(q.dtype == 0 || q.dtype == 1, "batch: dtype must be 0 (FP64) or 1 (FP32)");
    if (q.kind == XKBLAS_BATCH_GEMM) {
      if (q.m <= 0 || q.n <= 0) continue;
      const int tmx = (q.m + 63) / 64, tnx = (q.n + 63) / 64;
      for (int a = 0; a < tmx; ++a)
        for (int b = 0; b < tnx; ++b) {
          Cost c;
          c.w = BatchItem{i, a, b, 0};
          // 64 x 64 tile: k-loop work plus a fixed overhead
          c.c = 2.0 * 64 * 64 * std::max(q.k, 1) + 64 * 64 * 8.0;
          work.push_back(c);
        }
      pl->flops += 2.0 * q.m * q.n * (double)q.k;
      pl->bytes += elt * ((double)q.m * q.k + (double)q.k * q.n + 2.0 * q.m * q.n);
    } else {
      if (q.m <= 0 || q.n <= 0) continue;
      const int np = (q.n + 63) / 64;
      for (int b = 0; b < np; ++b) {
        Cost c;
        c.w = BatchItem{i, 0, b, 0};
        // left-looking blocks: one tile of k = 64 kb each plus the block's
        // solve, priced like a tile of k = 128
        const double nbk = (q.m + 63) / 64;
        c.c = 2.0 * 64 * 64 * 64 * nbk * (nbk - 1) / 2 + nbk * 2.0 * 64 * 64 * 128;
        work.push_back(c);
      }
      pl->flops += (double)q.m * q.m * q.n;
      pl->bytes += elt * (0.5 * (double)q.m * q.m + 2.0 * q.m * q.n);
    }
  }
  std::stable_sort(work.begin(), work.end(), [](const Cost& a, const Cost& b) { return a.c > b.c; });
  std::vector<BatchItem> items(work.size());
  for (size_t i = 0; i < work.size(); ++i) items[i] = work[i].w;
  pl->nitems = (int)items.size();
  pl->nprobs = count;
  XK_HIP(hipMalloc(&pl->d_probs, sizeof(BatchProblem) * std::max(count, 1)));
  XK_HIP(hipMalloc(&pl->d_items, sizeof(BatchItem) * std::max((int)items.size(), 1)));
  XK_HIP(hipMalloc(&pl->d_counter, sizeof(int)));
  if (count) XK_HIP(hipMemcpy(pl->d_probs, hp.data(), sizeof(BatchProblem) * count, hipMemcpyHostToDevice));
  if (!items.empty())
    XK_HIP(hipMemcpy(pl->d_items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice));
  // as many workgroups as can be resident at once
  int dev = 0, cus = 0, per_cu = 0;
  XK_HIP(hipGetDevice(&dev));
  XK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  XK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel, 256, 0));
  pl->grid = std::max(1, std::min(cus * std::max(per_cu, 1), std::max(pl->nitems, 1)));
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
