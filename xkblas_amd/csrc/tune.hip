// tune.hip — timing service of the autotuner (no reference counterpart).
//
// xkblas_tune_time runs one BLAS-3 call of a given shape and configuration
// on device-resident operands (a grow-only arena filled with seeded values
// by a kernel, so no host traffic is timed) and returns the kernel time
// measured with HIP events on the compute stream.  The configuration knobs
// are exactly the ones the tile-cost model chooses between: the FP64/FP32
// GEMM macro tile, the block order NB of TRSM/POTRF, the split count and the
// BF16 residue GEMM variant.
#include "blas3.h"
#include "kernels.h"
#include "runtime.h"
#include "split.h"
#include "../../include/xkblas.h"
#include <algorithm>
#include <climits>
#include <vector>

namespace xk {

// x -> uniform[-0.5, 0.5) from a 64-bit hash of (seed, index); diag > 0
// adds `diag` on the diagonal of an n x n matrix (SPD / dominant operands).
template <typename T>
__global__ void tune_fill_kernel(T* X, long ld, int m, int n, unsigned long long seed, double diag, int lower_only)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  unsigned long long z = seed ^ ((unsigned long long)i * 0x9E3779B97F4A7C15ull) ^ ((unsigned long long)j << 32);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  double v = (double)(z >> 11) / 9007199254740992.0 - 0.5;
  if (lower_only && i < j) v = 0.0;
  if (i == j) v += diag;
  X[i + (long)j * ld] = (T)v;
}

template <typename T>
static void tune_fill(hipStream_t s, T* X, long ld, int m, int n, unsigned long long seed, double diag, int lower)
{
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = std::min(65535, n - j0);
    hipLaunchKernelGGL((tune_fill_kernel<T>), dim3((m + 255) / 256, nn), dim3(256), 0, s, X + (long)j0 * ld, ld, m,
                       nn, seed + j0, diag, lower);
  }
  XK_HIP(hipGetLastError());
}

static void* g_arena[3] = {};
static size_t g_arena_size[3] = {};

static void* arena(int i, size_t bytes)
{
  if (g_arena_size[i] < bytes) {
    if (g_arena[i]) {
      XK_HIP(hipDeviceSynchronize());
      XK_HIP(hipFree(g_arena[i]));
    }
    XK_HIP(hipMalloc(&g_arena[i], bytes));
    g_arena_size[i] = bytes;
  }
  return g_arena[i];
}

template <typename T>
static double tune_run(int op, int mode, int ta, int tb, int m, int n, int k, int nb, int iters)
{
  Runtime& r = rt();
  Exec x = r.exec();
  x.mode = mode;
  hipStream_t s = x.s;
  const long lda = std::max(op == 0 ? (ta == NoTrans ? m : k) : (op == 1 ? n : m), 1);
  const long acols = op == 0 ? (ta == NoTrans ? k : m) : (op == 1 ? k : m);
  const long ldb = std::max(op == 0 ? (tb == NoTrans ? k : n) : m, 1);
  const long bcols = op == 0 ? (tb == NoTrans ? n : k) : n;
  const long ldc = std::max(op == 1 ? n : m, 1);
  const long ccols = n;
  T* A = (T*)arena(0, sizeof(T) * lda * std::max(acols, 1L));
  T* B = (T*)arena(1, sizeof(T) * ldb * std::max(bcols, 1L));
  T* C = (T*)arena(2, sizeof(T) * ldc * std::max(ccols, 1L));
  int* info = r.d_info + Runtime::INFO_SLOTS - 1;
  auto prepare = [&](int it) {
    if (op == 0) {
      tune_fill<T>(s, A, lda, (int)lda, (int)acols, 11 + it, 0.0, 0);
      tune_fill<T>(s, B, ldb, (int)ldb, (int)bcols, 23 + it, 0.0, 0);
      tune_fill<T>(s, C, ldc, (int)ldc, (int)ccols, 37 + it, 0.0, 0);
    } else if (op == 1) {
      tune_fill<T>(s, A, lda, (int)lda, (int)acols, 11 + it, 0.0, 0);
      tune_fill<T>(s, C, ldc, n, n, 37 + it, 0.0, 0);
    } else if (op == 2) {
      tune_fill<T>(s, A, lda, m, m, 11 + it, (double)m, 1);
      tune_fill<T>(s, B, ldb, m, n, 23 + it, 0.0, 0);
    } else {
      tune_fill<T>(s, A, lda, m, m, 11 + it, (double)m, 0);
      fill_int(s, info, 1, INT_MAX);
    }
  };
  hipEvent_t e0, e1;
  XK_HIP(hipEventCreate(&e0));
  XK_HIP(hipEventCreate(&e1));
  double best = 1e30;
  for (int it = 0; it < iters; ++it) {
    prepare(it);
    XK_HIP(hipEventRecord(e0, s));
    if (op == 0) dev_gemm<T>(x, ta, tb, m, n, k, T(1.5), A, lda, B, ldb, T(-0.5), C, ldc);
    else if (op == 1) dev_syrk<T>(x, Lower, NoTrans, n, k, T(1), A, lda, T(1), C, ldc);
    else if (op == 2) dev_trsm<T>(x, Left, Lower, NoTrans, NonUnit, m, n, T(1), A, lda, B, ldb, nb);
    else dev_potrf<T>(x, Lower, m, A, lda, nb, info);
    XK_HIP(hipEventRecord(e1, s));
    XK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    XK_HIP(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, (double)ms);
  }
  XK_HIP(hipEventDestroy(e0));
  XK_HIP(hipEventDestroy(e1));
  return best;
}

}  // namespace xk

using namespace xk;

extern "C" {

// op: 0 GEMM (C m x n, inner k, transA/transB), 1 SYRK (lower, NoTrans, C n x n,
// inner k), 2 TRSM (Left, Lower, NoTrans, Non-unit, B m x n, block order nb),
// 3 POTRF (lower, order m, block order nb).  dtype 0 FP64, 1 FP32; mode is
// the math mode (split modes apply to FP64).  tile: GEMM macro tile (0 auto,
// 32 (FP64; FP32 runs 64), 64, 128); split_count: 0 auto; split_kernel: 0 / 1.  Returns the best of
// `iters` kernel times in ms, or a negative value on bad arguments.
double xkblas_tune_time(int op, int dtype, int mode, int transA, int transB, int m, int n, int k, int tile, int nb,
                        int split_count, int split_kernel, int iters)
{
  if (op < 0 || op > 3 || m < 1 || n < 1 || k < 0 || iters < 1) return -1.0;
  if (dtype != 0 && dtype != 1) return -1.0;
  if (tile != 0 && tile != 32 && tile != 64 && tile != 128) return -1.0;
  xkblas_init();
  std::lock_guard<std::mutex> lock(api_mutex());
  Runtime& r = rt();
  XK_HIP(hipDeviceSynchronize());
  const int old_tile = gemm_get_tile(), old_split = r.nsplit, old_kernel = split_get_kernel();
  gemm_set_tile(tile);
  r.nsplit = split_count;
  split_set_kernel(split_kernel);
  if (op == 3) n = k = m;
  const int nbv = nb > 0 ? nb : 512;
  double ms = dtype == 0 ? tune_run<double>(op, mode, transA, transB, m, n, k, nbv, iters)
                         : tune_run<float>(op, 0, transA, transB, m, n, k, nbv, iters);
  gemm_set_tile(old_tile);
  r.nsplit = old_split;
  split_set_kernel(old_kernel);
  return ms;
}

// Knobs the tuner applies to subsequent calls.
void xkblas_set_gemm_tile(int tile)
{
  std::lock_guard<std::mutex> lock(api_mutex());
  gemm_set_tile(tile == 32 || tile == 64 || tile == 128 ? tile : 0);
}

void xkblas_set_split_kernel(int variant)
{
  std::lock_guard<std::mutex> lock(api_mutex());
  split_set_kernel(variant == 1 ? 1 : 0);
}

void xkblas_set_potrf_pairing(int mode)
{
  std::lock_guard<std::mutex> lock(api_mutex());
  potrf_set_pairing(mode);
}

int xkblas_split_side_bits(int k, int nsplit) { return split_side_bits(k, nsplit, false); }

int xkblas_split_info(int mode, int k, int nsplit, int* nmod, int* nplane)
{
  if (mode != XKBLAS_SPLIT_BF16_MATH && mode != XKBLAS_SPLIT_FP8_MATH) return -1;
  const bool fp8 = mode == XKBLAS_SPLIT_FP8_MATH;
  if (nmod) *nmod = split_choose_count(k, nsplit, fp8);
  if (nplane) *nplane = split_plane_count(k, nsplit, fp8);
  return split_side_bits(k, nsplit, fp8);
}

void xkblas_tune_release(void)
{
  std::lock_guard<std::mutex> lock(api_mutex());
  (void)hipDeviceSynchronize();
  for (int i = 0; i < 3; ++i) {
    if (g_arena[i]) (void)hipFree(g_arena[i]);
    g_arena[i] = nullptr;
    g_arena_size[i] = 0;
  }
}

}  // extern "C"
