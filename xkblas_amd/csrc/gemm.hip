// gemm.hip — FP64/FP32 MFMA GEMM for gfx950 (the tile body of every BLAS-3
// call in the library).
//
// Work decomposition.  One workgroup (4 waves, 256 lanes) owns a BM x BN
// macro tile of C; each wave owns a (BM/2) x (BN/2) quarter as (BM/32) x
// (BN/32) accumulators of 16x16 (v_mfma_*_16x16x4: 4 doubles / floats per
// lane).  The k loop streams BK-deep slices of op(A) and op(B) through two
// LDS buffers: the next slice is loaded from HBM into registers while the
// current one feeds the MFMAs, then written to the other buffer; one barrier
// per slice.  LDS rows are k-major (op(A)[k][m], op(B)[k][n]) so every MFMA
// operand read is 16 consecutive elements per 16 lanes.
//
// The MFMA is issued with op(B) as its first operand, so the accumulator holds
// C^T in the hardware's output layout: lane l, register r holds
// C[m = 16*tile + l%16][n = 16*tile + q(l,r)] with q = l/16 + 4r for
// v_mfma_f64_16x16x4f64 and q = 4(l/16) + r for v_mfma_f32_16x16x4f32.
// Sixteen consecutive lanes then store 16 consecutive rows of one column:
// 128-byte segments of the column-major C.
//
// Tile order.  Workgroup ids are dealt round-robin to the 8 XCDs; the id is
// remapped so each XCD receives a contiguous range of tiles, and tiles are
// rastered in groups of 8 tile rows so that concurrently running tiles of one
// XCD share op(A) panels in its L2.
#include "common.h"
#include "tiles.h"
#include "kernels.h"

namespace xk {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct AccT;
template <> struct AccT<double> { typedef d4_t type; };
template <> struct AccT<float> { typedef f4_t type; };

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
struct GemmParams {
  int m, n, k;
  int tiles_m, tiles_n;
  int tri;
  int kmode;       // inner range per tile for a triangular operand (KRANGE_*)
  int vecA, vecB;  // 16-byte vector loads allowed (base and ld aligned)
  int prio;        // raise the waves' issue priority (latency-bound chain steps)
  T alpha, beta;
  const T* A;
  long lda, sA;
  const T* B;
  long ldb, sB;
  T* C;
  long ldc, sC;
};

// Load V consecutive elements starting at p (16-byte aligned when vec).
template <typename T, int V>
__device__ __forceinline__ void load_run(const T* p, T (&r)[V])
{
  static_assert((V * sizeof(T)) % 16 == 0, "run must be a multiple of 16 bytes");
  typedef float4 v16;
  const v16* q = reinterpret_cast<const v16*>(p);
#pragma unroll
  for (int i = 0; i < (int)(V * sizeof(T) / 16); ++i) {
    v16 x = q[i];
    __builtin_memcpy(reinterpret_cast<char*>(r) + 16 * i, &x, 16);
  }
}

template <typename T, int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmParams<T> p)
{
  constexpr int PAD = 16 / sizeof(T);
  constexpr int LA = BM + PAD, LB = BN + PAD;
  constexpr int VA = BM * BK / 256;  // elements of op(A) per lane per slice
  constexpr int VB = BN * BK / 256;
  constexpr int WTM = BM / 2, WTN = BN / 2;  // wave tile
  constexpr int TM = WTM / 16, TN = WTN / 16;
  typedef typename AccT<T>::type acc_t;

  __shared__ T As[2][BK][LA];
  __shared__ T Bs[2][BK][LB];
  if (p.prio) __builtin_amdgcn_s_setprio(3);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  const T* __restrict__ A = p.A + (long)blockIdx.z * p.sA;
  const T* __restrict__ B = p.B + (long)blockIdx.z * p.sB;
  T* __restrict__ C = p.C + (long)blockIdx.z * p.sC;

  // ---- tile of this workgroup
  int tm, tn;
  {
    int pid = blockIdx.x;
    const int total = gridDim.x;
    if (p.tri == TRI_FULL) {
      // a limited inner range makes the work per tile grow along the tile
      // rows or columns: no XCD blocking then (consecutive ids go to
      // different XCDs, which then share light and heavy tiles alike), and
      // the heaviest tiles are dispatched first
      if (p.kmode == KRANGE_ALL && (total & 7) == 0) pid = (pid & 7) * (total >> 3) + (pid >> 3);
      constexpr int GM = 8;
      const int group = GM * p.tiles_n;
      const int g = pid / group;
      const int first = g * GM;
      const int gm = min(p.tiles_m - first, GM);
      const int r = pid - g * group;
      tm = first + r % gm;
      tn = r / gm;
      if (p.kmode == KRANGE_A_LOWER) tm = p.tiles_m - 1 - tm;
      else if (p.kmode == KRANGE_B_UPPER) tn = p.tiles_n - 1 - tn;
    } else {
      int i, j;
      tri_tile<8>(pid, total, p.tiles_m, i, j);
      if (p.tri == TRI_LOWER) { tm = i; tn = j; }
      else { tm = j; tn = i; }
    }
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int M = p.m, N = p.n, K = p.k;

  // ---- per-lane load coordinates
  // op(A) slice BM x BK.  !TA: runs along m.  TA: runs along k.
  int a_r, a_c;  // row (m) and column (k) of the run start inside the slice
  if (!TA) { a_c = tid / (BM / VA); a_r = (tid % (BM / VA)) * VA; }
  else     { a_r = tid / (BK / VA); a_c = (tid % (BK / VA)) * VA; }
  // op(B) slice BK x BN.  !TB: runs along k.  TB: runs along n.
  int b_r, b_c;  // k and n of the run start
  if (!TB) { b_c = tid / (BK / VB); b_r = (tid % (BK / VB)) * VB; }
  else     { b_r = tid / (BN / VB); b_c = (tid % (BN / VB)) * VB; }

  T ra[VA], rb[VB];

  auto load_slice = [&](int k0) {
    const bool fullA = p.vecA && (m0 + BM <= M) && (k0 + BK <= K);
    const bool fullB = p.vecB && (n0 + BN <= N) && (k0 + BK <= K);
    if (!TA) {
      const T* src = A + (long)(m0 + a_r) + (long)(k0 + a_c) * p.lda;
      if (fullA) load_run<T, VA>(src, ra);
      else {
#pragma unroll
        for (int v = 0; v < VA; ++v)
          ra[v] = (m0 + a_r + v < M && k0 + a_c < K) ? src[v] : T(0);
      }
    } else {
      const T* src = A + (long)(k0 + a_c) + (long)(m0 + a_r) * p.lda;
      if (fullA) load_run<T, VA>(src, ra);
      else {
#pragma unroll
        for (int v = 0; v < VA; ++v)
          ra[v] = (m0 + a_r < M && k0 + a_c + v < K) ? src[v] : T(0);
      }
    }
    if (!TB) {
      const T* src = B + (long)(k0 + b_r) + (long)(n0 + b_c) * p.ldb;
      if (fullB) load_run<T, VB>(src, rb);
      else {
#pragma unroll
        for (int v = 0; v < VB; ++v)
          rb[v] = (k0 + b_r + v < K && n0 + b_c < N) ? src[v] : T(0);
      }
    } else {
      const T* src = B + (long)(n0 + b_c) + (long)(k0 + b_r) * p.ldb;
      if (fullB) load_run<T, VB>(src, rb);
      else {
#pragma unroll
        for (int v = 0; v < VB; ++v)
          rb[v] = (k0 + b_r < K && n0 + b_c + v < N) ? src[v] : T(0);
      }
    }
  };

  auto store_slice = [&](int buf) {
    if (!TA) {
#pragma unroll
      for (int v = 0; v < VA; ++v) As[buf][a_c][a_r + v] = ra[v];
    } else {
#pragma unroll
      for (int v = 0; v < VA; ++v) As[buf][a_c + v][a_r] = ra[v];
    }
    if (!TB) {
#pragma unroll
      for (int v = 0; v < VB; ++v) Bs[buf][b_r + v][b_c] = rb[v];
    } else {
#pragma unroll
      for (int v = 0; v < VB; ++v) Bs[buf][b_r][b_c + v] = rb[v];
    }
  };

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // inner range: all of k, or the part where a triangular op(A) / op(B) of
  // this tile row / column is not zero (the rest would add exact zeros)
  int kbeg = 0, kend = K;
  if (p.kmode == KRANGE_A_LOWER) kend = min(K, m0 + BM);
  else if (p.kmode == KRANGE_A_UPPER) kbeg = m0;
  else if (p.kmode == KRANGE_B_UPPER) kend = min(K, n0 + BN);
  else if (p.kmode == KRANGE_B_LOWER) kbeg = n0;
  const int kt0 = kbeg / BK;
  const int nk = (kend + BK - 1) / BK;
  const int lr = lane & 15, lk = lane >> 4;
  if (nk > kt0) {
    load_slice(kt0 * BK);
    store_slice(0);
    __syncthreads();
    int buf = 0;
    for (int kt = kt0; kt < nk; ++kt) {
      const bool more = kt + 1 < nk;
      if (more) load_slice((kt + 1) * BK);
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        T af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[buf][ks * 4 + lk][wm * WTM + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = Bs[buf][ks * 4 + lk][wn * WTN + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma16(bf[j], af[i], acc[i][j]);
      }
      if (more) store_slice(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }

  // ---- epilogue: C = alpha acc + beta C on the selected triangle
  const T alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int gm = m0 + wm * WTM + i * 16 + lr;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // output row of the 16x16 MFMA block held by (lane, r): f64 uses
        // lane/16 + 4r, f32 uses 4(lane/16) + r (measured on gfx950)
        const int gn = n0 + wn * WTN + j * 16 + (sizeof(T) == 8 ? lk + 4 * r : lk * 4 + r);
        bool ok = gm < M && gn < N;
        if (p.tri == TRI_LOWER) ok = ok && gm >= gn;
        else if (p.tri == TRI_UPPER) ok = ok && gm <= gn;
        if (ok) {
          T* c = C + (long)gm + (long)gn * p.ldc;
          T v = alpha * acc[i][j][r];
          if (beta != T(0)) v += beta * *c;
          *c = v;
        }
      }
    }
  }
}

static int g_force_tile = 0;
void gemm_set_tile(int bm) { g_force_tile = bm; }

// Products launched on this stream run their waves at raised issue priority
// (the latency-bound diagonal chain of POTRF, beside full-chip updates).
static hipStream_t g_prio_stream = nullptr;
void gemm_set_priority_stream(hipStream_t s) { g_prio_stream = s; }
int gemm_get_tile() { return g_force_tile; }

template <typename T, int BM, int BK>
static void launch_shape(hipStream_t s, bool ta, bool tb, GemmParams<T>& p, int batch)
{
  p.tiles_m = (p.m + BM - 1) / BM;
  p.tiles_n = (p.n + BM - 1) / BM;
  long ntiles = (p.tri == TRI_FULL) ? (long)p.tiles_m * p.tiles_n
                                    : (long)p.tiles_m * (p.tiles_m + 1) / 2;
  XK_CHECK(ntiles < (1L << 31), "gemm grid too large");
  dim3 grid((unsigned)ntiles, 1, (unsigned)batch), block(256);
  if (!ta && !tb) hipLaunchKernelGGL((gemm_kernel<T, BM, BM, BK, false, false>), grid, block, 0, s, p);
  else if (!ta && tb) hipLaunchKernelGGL((gemm_kernel<T, BM, BM, BK, false, true>), grid, block, 0, s, p);
  else if (ta && !tb) hipLaunchKernelGGL((gemm_kernel<T, BM, BM, BK, true, false>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((gemm_kernel<T, BM, BM, BK, true, true>), grid, block, 0, s, p);
  XK_HIP(hipGetLastError());
}

template <typename T>
void gemm(hipStream_t s, int transA, int transB, int m, int n, int k, T alpha, const T* A, long lda,
          const T* B, long ldb, T beta, T* C, long ldc, int tri, int batch, long sA, long sB, long sC, int kmode)
{
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (tri != TRI_FULL) XK_CHECK(m == n, "triangular gemm output must be square");
  GemmParams<T> p;
  p.m = m; p.n = n; p.k = k; p.tri = tri; p.kmode = kmode;
  p.alpha = alpha; p.beta = beta;
  p.A = A; p.lda = lda; p.sA = sA;
  p.B = B; p.ldb = ldb; p.sB = sB;
  p.C = C; p.ldc = ldc; p.sC = sC;
  auto aligned = [](const void* q, long ld, long stride) {
    return ((uintptr_t)q % 16 == 0) && ((ld * (long)sizeof(T)) % 16 == 0) && ((stride * (long)sizeof(T)) % 16 == 0);
  };
  p.prio = s == g_prio_stream && s != nullptr;
  p.vecA = aligned(A, lda, sA);
  p.vecB = aligned(B, ldb, sB);
  const bool ta = transA != NoTrans, tb = transB != NoTrans;
  int bm = g_force_tile;
  if (bm == 0) {
    // FP64: 64 x 64 tiles at every size (measured on MI355X: 55.8 vs 47.1
    // TFLOP/s at 16384^3, more resident waves per CU); FP32: 128 x 128 once
    // there are enough tiles to fill the chip
    long t128 = (long)((m + 127) / 128) * ((n + 127) / 128) * batch;
    bm = (sizeof(T) == 8 || m <= 64 || n <= 64 || t128 < 160) ? 64 : 128;
    // FP64 products too small to give every CU two 64 x 64 tiles (the
    // latency-bound steps of TRSM sweeps and diagonal-block chains): 32 x 32
    // tiles, four times the workgroups; the k order per element, hence the
    // result, is the same
    long t64 = (long)((m + 63) / 64) * ((n + 63) / 64);
    if (p.tri != TRI_FULL) t64 = t64 / 2 + (m + 63) / 64;
    if (sizeof(T) == 8 && t64 * batch < 512) bm = 32;
  }
  if (bm == 128) launch_shape<T, 128, 16>(s, ta, tb, p, batch);
  else if (bm == 32) {
    if constexpr (sizeof(T) == 8) launch_shape<T, 32, 16>(s, ta, tb, p, batch);
    else launch_shape<T, 64, 16>(s, ta, tb, p, batch);
  } else launch_shape<T, 64, 16>(s, ta, tb, p, batch);
}

template void gemm<double>(hipStream_t, int, int, int, int, int, double, const double*, long, const double*, long,
                           double, double*, long, int, int, long, long, long, int);
template void gemm<float>(hipStream_t, int, int, int, int, int, float, const float*, long, const float*, long,
                          float, float*, long, int, int, long, long, long, int);

}  // namespace xk
