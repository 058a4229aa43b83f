// refine.hip — element-wise kernels around the GEMM: expansion of one stored
// triangle to the full symmetric matrix (SYMM, the SPD solve's residual) or to
// a full triangular one (TRMM), column max-norms and absolute sums, X += D.
// All are HBM-bound single passes with coalesced column-major accesses.
#include "kernels.h"
#include <algorithm>

namespace xk {

// Y (and Z when non-null) <- the symmetric matrix whose uplo triangle is
// stored in A, tile columns jt0 .. jt0 + gridDim.y - 1.  One 64 x 64 tile per
// workgroup (256 threads as 64 rows x 4 column phases); a tile on the
// mirrored side reads its transposed source tile coalesced and transposes it
// through LDS.
template <typename T>
__global__ __launch_bounds__(256) void symm_expand_kernel(int uplo, int n, int jt0, const T* __restrict__ A,
                                                          long lda, T* __restrict__ Y, long ldy, T* __restrict__ Z,
                                                          long ldz)
{
  __shared__ T t[64][65];
  const int bi = blockIdx.x, bj = jt0 + blockIdx.y;
  const int i0 = bi * 64, j0 = bj * 64;
  const bool lower = uplo == Lower;
  const bool mirror = lower ? bi < bj : bi > bj;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i = i0 + tx;
  if (!mirror) {
    const bool diag = bi == bj;
    for (int c = ty; c < 64; c += 4) {
      const int j = j0 + c;
      if (i < n && j < n) {
        const bool stored = !diag || (lower ? i >= j : i <= j);
        const T v = stored ? A[i + (long)j * lda] : A[j + (long)i * lda];
        Y[i + (long)j * ldy] = v;
        if (Z) Z[i + (long)j * ldz] = v;
      }
    }
    return;
  }
  // t[c][r] = A(j0 + r, i0 + c)
  for (int c = ty; c < 64; c += 4) {
    const int r = j0 + tx, col = i0 + c;
    if (r < n && col < n) t[c][tx] = A[r + (long)col * lda];
  }
  __syncthreads();
  for (int c = ty; c < 64; c += 4) {
    const int j = j0 + c;
    if (i < n && j < n) {
      const T v = t[tx][c];  // A(j, i)
      Y[i + (long)j * ldy] = v;
      if (Z) Z[i + (long)j * ldz] = v;
    }
  }
}

// Y <- the n x n triangular matrix whose uplo triangle is stored in A: zeros
// outside the triangle, ones on the diagonal when unit.
template <typename T>
__global__ __launch_bounds__(256) void tri_expand_kernel(int uplo, int unit, int n, int j0, const T* __restrict__ A,
                                                         long lda, T* __restrict__ Y, long ldy)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= n || j >= n) return;
  const bool inside = uplo == Lower ? i > j : i < j;
  Y[i + (long)j * ldy] = i == j ? (unit ? T(1) : A[i + (long)j * lda]) : (inside ? A[i + (long)j * lda] : T(0));
}

// NaN-propagating max of two non-negative values.
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || a > b) ? a : b; }

// out[j] = max_i |X(i, j)| (op 0) or sum_i |X(i, j)| (op 1); one workgroup per
// column, 256 threads striding down the column.
__global__ __launch_bounds__(256) void col_reduce_kernel(int op, int m, const double* __restrict__ X, long ldx,
                                                         double* __restrict__ out)
{
  __shared__ double red[256];
  const int j = blockIdx.x;
  const double* col = X + (long)j * ldx;
  double v = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) {
    const double a = fabs(col[i]);
    v = op ? v + a : nmax(v, a);
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = op ? red[threadIdx.x] + red[threadIdx.x + w]
                                                    : nmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

__global__ __launch_bounds__(256) void add2d_kernel(int m, int n, const double* __restrict__ D, long ldd,
                                                    double* __restrict__ X, long ldx)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i < m && j < n) X[i + (long)j * ldx] += D[i + (long)j * ldd];
}

// ---- accurate residual R = B - A X for refinement
// A FP64 GEMM sums its k products in one chain per element, so its rounding
// error grows with k (typically sqrt(k) u |A| |X|) and refinement stalls at
// that level.  Here the k products of each element are summed by MFMA in
// runs of RES_SUB = 64 only; the run sums are added without error
// (TwoSum, the rounding errors kept in a second word) per workgroup over a
// k-range of RES_KR (XKBLAS_RES_KR, a multiple of 64), and the ranges'
// (sum, error) pairs are added to B in the same way by the reduction kernel.
// The remaining error is that of the 64-term runs and one final rounding:
// about sqrt(k / 64) times smaller.  RES_KR = 4096 rather than 1024: a
// quarter of the partial pairs written and read back (N = 32768, 64 right-hand
// sides: 3.65 + 0.19 -> 3.53 + 0.05 ms per residual).
constexpr int RES_SUB = 64, RES_KR = 4096;
constexpr int RES_WAVES = 8;           // 16 rows each: 128 rows per workgroup
constexpr int RES_LDK = RES_SUB + 4;   // padded k stride of the staged X run
typedef double rd4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
  s = a + b;
  const double bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

// Partial sums over the k-range of blockIdx.z: P[z * sp + i + j ldp] =
// (sum, error) of sum_k A(i, k) X(k, j).  Workgroup: 128 rows x 64 columns,
// wave w the rows 16 w .. 16 w + 15, four 16 x 16 blocks of
// v_mfma_f64_16x16x4f64.  Each 64-k run of X is staged once in LDS for the
// eight waves (Xs[j * RES_LDK + k]); A is read once, straight from memory;
// the loads of the next run (A and X, in registers) are issued before the
// products of the current one.
__global__ __launch_bounds__(512) void resid_partial_kernel(int n, int nrhs, const double* __restrict__ A, long lda,
                                                            const double* __restrict__ X, long ldx,
                                                            double2* __restrict__ P, long ldp, long sp, int kr)
{
  __shared__ double Xs[64 * RES_LDK];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = blockIdx.x * (16 * RES_WAVES) + wave * 16, j0 = blockIdx.y * 64;
  const int k0 = blockIdx.z * kr, k1 = min(n, k0 + kr);
  const int ia = i0 + li;
  const bool rowok = ia < n;
  double hi[4][4], lo[4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) hi[b][r] = lo[b][r] = 0.0;
  constexpr int NX = 64 * RES_SUB / (64 * RES_WAVES), NA = RES_SUB / 4;
  double xv[NX], av[NA], an[NA];
  // the loads of run ks (X into xv, A into the given registers)
  auto load_run = [&](int ks, double* a) {
    const int ke = min(k1, ks + RES_SUB);
#pragma unroll
    for (int u = 0; u < NX; ++u) {
      const int e = tid + 64 * RES_WAVES * u, j = e >> 6, kk = ks + (e & 63);
      xv[u] = (kk < ke && j0 + j < nrhs) ? X[kk + (long)(j0 + j) * ldx] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      const int kk = ks + 4 * t + lk;
      a[t] = (rowok && kk < ke) ? A[ia + (long)kk * lda] : 0.0;
    }
  };
  load_run(k0, an);
  for (int ks = k0; ks < k1; ks += RES_SUB) {
#pragma unroll
    for (int t = 0; t < NA; ++t) av[t] = an[t];
    __syncthreads();  // the previous run's X reads are done
#pragma unroll
    for (int u = 0; u < NX; ++u) {
      const int e = tid + 64 * RES_WAVES * u;
      Xs[(e >> 6) * RES_LDK + (e & 63)] = xv[u];
    }
    __syncthreads();
    // the next run's loads fly under this run's products
    if (ks + RES_SUB < k1) load_run(ks + RES_SUB, an);
    rd4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < RES_SUB / 4; ++t) {
      // (x first: lane holds element (row li, column lk + 4 r) of the block)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[(16 * b + li) * RES_LDK + 4 * t + lk], av[t], acc[b], 0, 0,
                                                      0);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double t, e;
        two_sum(hi[b][r], acc[b][r], t, e);
        hi[b][r] = t;
        lo[b][r] += e;
      }
  }
  if (!rowok) return;
  double2* out = P + blockIdx.z * sp;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = j0 + 16 * b + lk + 4 * r;
      if (j < nrhs) out[ia + (long)j * ldp] = make_double2(hi[b][r], lo[b][r]);
    }
}

// R(i, j) = B(i, j) - sum_z P[z](i, j), the pairs added without error and
// the errors summed apart.
__global__ __launch_bounds__(256) void resid_reduce_kernel(int n, int nz, const double* __restrict__ B, long ldb,
                                                           const double2* __restrict__ P, long ldp, long sp,
                                                           double* __restrict__ R, long ldr)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  double s = B[i + (long)j * ldb], c = 0.0;
  for (int z = 0; z < nz; ++z) {
    const double2 q = P[z * sp + i + (long)j * ldp];
    double t, e;
    two_sum(s, -q.x, t, e);
    s = t;
    c += e - q.y;
  }
  R[i + (long)j * ldr] = s + c;
}

static int res_kr()
{
  static const int v = getenv("XKBLAS_RES_KR") ? atoi(getenv("XKBLAS_RES_KR")) / RES_SUB * RES_SUB : RES_KR;
  return v > 0 ? v : RES_KR;
}

size_t resid_scratch_bytes(int n, int nrhs)
{
  const long ldp = (n + 63) / 64 * 64;
  const long nz = (n + res_kr() - 1) / res_kr();
  return sizeof(double2) * (size_t)nz * ldp * nrhs;
}

void resid_accurate(hipStream_t s, int n, int nrhs, const double* A, long lda, const double* X, long ldx,
                    const double* B, long ldb, double* R, long ldr, void* work)
{
  if (n <= 0 || nrhs <= 0) return;
  const long ldp = (n + 63) / 64 * 64;
  const int nz = (n + res_kr() - 1) / res_kr();
  const long sp = ldp * nrhs;
  double2* P = static_cast<double2*>(work);
  XK_CHECK((nrhs + 63) / 64 <= 65535 && nz <= 65535 && nrhs <= 65535, "residual grid");
  hipLaunchKernelGGL(resid_partial_kernel, dim3((n + 16 * RES_WAVES - 1) / (16 * RES_WAVES), (nrhs + 63) / 64, nz),
                     dim3(64 * RES_WAVES), 0, s, n, nrhs, A, lda, X, ldx, P, ldp, sp, res_kr());
  hipLaunchKernelGGL(resid_reduce_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, nz, B, ldb, P, ldp, sp,
                     R, ldr);
  XK_HIP(hipGetLastError());
}

template <typename T>
void symm_expand(hipStream_t s, int uplo, int n, const T* A, long lda, T* Y, long ldy, T* Z, long ldz, int c0,
                 int c1)
{
  if (n <= 0) return;
  if (c1 < 0 || c1 > n) c1 = n;
  if (c0 >= c1) return;  // empty range (c0 may then be n, of any value)
  XK_CHECK(c0 >= 0 && c0 % 64 == 0, "symm_expand: first column must be a multiple of 64");
  const int nt = (n + 63) / 64, jt0 = c0 / 64, ntc = (c1 + 63) / 64 - jt0;
  XK_CHECK(nt < 65536, "symm_expand: order too large");
  hipLaunchKernelGGL((symm_expand_kernel<T>), dim3(nt, ntc), dim3(256), 0, s, uplo, n, jt0, A, lda, Y, ldy, Z, ldz);
  XK_HIP(hipGetLastError());
}
template void symm_expand<double>(hipStream_t, int, int, const double*, long, double*, long, double*, long, int,
                                  int);
template void symm_expand<float>(hipStream_t, int, int, const float*, long, float*, long, float*, long, int, int);

template <typename T>
void tri_expand(hipStream_t s, int uplo, int unit, int n, const T* A, long lda, T* Y, long ldy)
{
  if (n <= 0) return;
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = std::min(65535, n - j0);
    hipLaunchKernelGGL((tri_expand_kernel<T>), dim3((n + 255) / 256, nn), dim3(256), 0, s, uplo, unit, n, j0, A, lda,
                       Y, ldy);
  }
  XK_HIP(hipGetLastError());
}
template void tri_expand<double>(hipStream_t, int, int, int, const double*, long, double*, long);
template void tri_expand<float>(hipStream_t, int, int, int, const float*, long, float*, long);

void col_reduce(hipStream_t s, int op, int m, int n, const double* X, long ldx, double* out)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(col_reduce_kernel, dim3(n), dim3(256), 0, s, op, m, X, ldx, out);
  XK_HIP(hipGetLastError());
}

// FP32 images for the refinement's solves.  Z <- float(Y 2^-e_j) per column j,
// with e_j the exponent of colmax[j] (0 when it is 0 or not finite): the
// scaling is exact and keeps small residuals away from FP32's underflow.
__global__ __launch_bounds__(256) void to_f32_kernel(int m, int n, const double* __restrict__ Y, long ldy,
                                                     const double* __restrict__ colmax, float* __restrict__ Z, long ldz)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  const double c = colmax ? colmax[j] : 0.0;
  const int e = (c > 0.0 && c <= 1.7976931348623157e308) ? __builtin_amdgcn_frexp_exp(c) : 0;
  Z[i + (long)j * ldz] = (float)__builtin_ldexp(Y[i + (long)j * ldy], -e);
}

// Y <- double(Z) 2^e_j (acc: Y += ...), the inverse of to_f32_kernel's scaling.
__global__ __launch_bounds__(256) void from_f32_kernel(int m, int n, const float* __restrict__ Z, long ldz,
                                                       const double* __restrict__ colmax, double* __restrict__ Y,
                                                       long ldy, int acc)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  const double c = colmax ? colmax[j] : 0.0;
  const int e = (c > 0.0 && c <= 1.7976931348623157e308) ? __builtin_amdgcn_frexp_exp(c) : 0;
  const double v = __builtin_ldexp((double)Z[i + (long)j * ldz], e);
  double* y = Y + i + (long)j * ldy;
  *y = acc ? *y + v : v;
}

void to_f32(hipStream_t s, int m, int n, const double* Y, long ldy, const double* colmax, float* Z, long ldz)
{
  if (m <= 0 || n <= 0) return;
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = std::min(65535, n - j0);
    hipLaunchKernelGGL(to_f32_kernel, dim3((m + 255) / 256, nn), dim3(256), 0, s, m, nn, Y + (long)j0 * ldy, ldy,
                       colmax ? colmax + j0 : nullptr, Z + (long)j0 * ldz, ldz);
  }
  XK_HIP(hipGetLastError());
}

void from_f32(hipStream_t s, int m, int n, const float* Z, long ldz, const double* colmax, double* Y, long ldy,
              bool acc)
{
  if (m <= 0 || n <= 0) return;
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = std::min(65535, n - j0);
    hipLaunchKernelGGL(from_f32_kernel, dim3((m + 255) / 256, nn), dim3(256), 0, s, m, nn, Z + (long)j0 * ldz, ldz,
                       colmax ? colmax + j0 : nullptr, Y + (long)j0 * ldy, ldy, acc ? 1 : 0);
  }
  XK_HIP(hipGetLastError());
}

void add2d(hipStream_t s, int m, int n, const double* D, long ldd, double* X, long ldx)
{
  if (m <= 0 || n <= 0) return;
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = std::min(65535, n - j0);
    hipLaunchKernelGGL(add2d_kernel, dim3((m + 255) / 256, nn), dim3(256), 0, s, m, nn, D + (long)j0 * ldd, ldd,
                       X + (long)j0 * ldx, ldx);
  }
  XK_HIP(hipGetLastError());
}

}  // namespace xk
