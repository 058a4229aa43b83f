// factor.hip — latency-bound diagonal-block kernels for gfx950:
// Cholesky + inverse of blocks of order <= 64, triangular inverse, and small
// helpers (2-D copy, triangle zeroing, int fill).
//
// A leaf is one wave (64 lanes) per block; the padded rows/columns beyond n
// are the identity, so every block is handled as a 64 x 64 problem.  The
// Cholesky leaf keeps the block in registers (one row per lane, cross-lane
// reads for the broadcasts) and writes the factor and its inverse in one
// launch; the triangular-inverse leaf holds the block in LDS (column stride
// 65) and solves one column of inv(L) per lane.
#include "common.h"
#include "kernels.h"
#include <climits>

namespace xk {

constexpr int LEAF = 64;
constexpr int LLD = LEAF + 1;  // padded LDS column stride

__device__ __forceinline__ double bcast(double x, int src)
{
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float bcast(float x, int src)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}

// Row `lane` of the lower-triangular view L of block A:
//   uplo == Lower: L = tril(A); uplo == Upper: L = triu(A)^T.
template <typename T>
__device__ __forceinline__ T lower_view(const T* A, long lda, int uplo, int n, int i, int j)
{
  if (i >= n || j >= n) return (i == j) ? T(1) : T(0);
  if (j > i) return T(0);
  return uplo == Lower ? A[i + (long)j * lda] : A[j + (long)i * lda];
}

// inv(L) for the 64 x 64 lower-triangular L held in LDS (column-major,
// stride LLD); lane j computes column j.  unit: diagonal taken as 1.
template <typename T>
__device__ __forceinline__ void lower_inverse_to(const T* Ls, int lane, bool unit, T* Tinv, long ldinv, int n,
                                                 int uplo)
{
  T x[LEAF];
#pragma unroll
  for (int i = 0; i < LEAF; ++i) {
    T s = (i == lane) ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < i; ++k) s -= Ls[i + k * LLD] * x[k];
    x[i] = unit ? s : s / Ls[i + i * LLD];
    if (i < lane) x[i] = T(0);
  }
  // column `lane` of inv(L): X[i][lane].  Lower: Tinv[i + lane*ld];
  // Upper (inv(U) = inv(L)^T): Tinv[lane + i*ld].
  if (lane < n) {
#pragma unroll
    for (int i = 0; i < LEAF; ++i) {
      if (i < n) {
        if (uplo == Lower) Tinv[i + (long)lane * ldinv] = x[i];
        else Tinv[lane + (long)i * ldinv] = x[i];
      }
    }
  }
}

// Cholesky factor and its inverse for one block of order <= 64, in
// registers: lane i holds row i of the lower-triangular view (x) and later
// column i of inv(L) (y); the values of other rows come from cross-lane
// reads (v_readlane), so there is no LDS traffic and no barrier.  Step j
// takes the pivot from lane j, scales column j and applies the rank-1 update
// x[k] -= L[i][j] L[k][j] for k > j (entries above the diagonal become
// garbage and are cleared before they are read).  The inverse is a forward
// substitution, column `lane` per lane.
template <typename T>
__global__ __launch_bounds__(64) void potrf_leaf_kernel(int uplo, int n, T* A, long lda, long sA, T* Tinv,
                                                        long ldinv, long sInv, int* info, int base)
{
  const int lane = threadIdx.x;
  A += (long)blockIdx.x * sA;
  Tinv += (long)blockIdx.x * sInv;
  T x[LEAF];
#pragma unroll
  for (int j = 0; j < LEAF; ++j) x[j] = lower_view(A, lda, uplo, n, lane, j);

  bool failed = false;
#pragma unroll
  for (int j = 0; j < LEAF; ++j) {
    const T d = bcast(x[j], j);
    if (!(d > T(0)) && !failed && j < n) {
      failed = true;
      if (lane == 0) atomicMin(info, base + j + 1);
    }
    const T s = sqrt(d);
    x[j] = lane > j ? x[j] / s : (lane == j ? s : T(0));
#pragma unroll
    for (int k = j + 1; k < LEAF; ++k) x[k] = fma(-x[j], bcast(x[j], k), x[k]);
  }
  // factor back to the stored triangle
#pragma unroll
  for (int j = 0; j < LEAF; ++j) {
    if (lane < n && j < n && j <= lane) {
      if (uplo == Lower) A[lane + (long)j * lda] = x[j];
      else A[j + (long)lane * lda] = x[j];
    }
  }
  if (!Tinv) return;
  // inv(L): lane j computes column j, y[i] = inv(L)[i][j]
  T y[LEAF];
#pragma unroll
  for (int i = 0; i < LEAF; ++i) {
    T acc = (i == lane) ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < i; ++k) acc = fma(-bcast(x[k], i), y[k], acc);
    y[i] = i < lane ? T(0) : acc / bcast(x[i], i);
  }
  if (lane < n) {
#pragma unroll
    for (int i = 0; i < LEAF; ++i) {
      if (i < n) {
        if (uplo == Lower) Tinv[i + (long)lane * ldinv] = y[i];
        else Tinv[lane + (long)i * ldinv] = y[i];
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(64) void trtri_leaf_kernel(int uplo, int diag, int n, const T* A, long lda, long sA,
                                                        T* Tinv, long ldinv, long sInv)
{
  __shared__ T Ls[LEAF * LLD];
  const int lane = threadIdx.x;
  A += (long)blockIdx.x * sA;
  Tinv += (long)blockIdx.x * sInv;
#pragma unroll 4
  for (int j = 0; j < LEAF; ++j) Ls[lane + j * LLD] = lower_view(A, lda, uplo, n, lane, j);
  __syncthreads();
  lower_inverse_to(Ls, lane, diag == Unit, Tinv, ldinv, n, uplo);
}

template <typename T>
void potrf_leaf(hipStream_t s, int uplo, int n, T* A, long lda, long sA, T* Tinv, long ldinv, long sInv,
                int batch, int* info, int base)
{
  XK_CHECK(n >= 1 && n <= LEAF, "potrf leaf order must be in [1,64]");
  if (batch <= 0) return;
  hipLaunchKernelGGL((potrf_leaf_kernel<T>), dim3(batch), dim3(64), 0, s, uplo, n, A, lda, sA, Tinv, ldinv, sInv,
                     info, base);
  XK_HIP(hipGetLastError());
}

template <typename T>
void trtri_leaf(hipStream_t s, int uplo, int diag, int n, const T* A, long lda, long sA, T* Tinv, long ldinv,
                long sInv, int batch)
{
  XK_CHECK(n >= 1 && n <= LEAF, "trtri leaf order must be in [1,64]");
  if (batch <= 0) return;
  hipLaunchKernelGGL((trtri_leaf_kernel<T>), dim3(batch), dim3(64), 0, s, uplo, diag, n, A, lda, sA, Tinv, ldinv,
                     sInv);
  XK_HIP(hipGetLastError());
}

template <typename T>
__global__ void copy2d_kernel(int m, int n, const T* X, long ldx, T* Y, long ldy)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < m && j < n) Y[i + (long)j * ldy] = X[i + (long)j * ldx];
}

template <typename T>
void copy2d(hipStream_t s, int m, int n, const T* X, long ldx, T* Y, long ldy)
{
  if (m <= 0 || n <= 0) return;
  XK_CHECK(n < 65536 * 16, "copy2d: too many columns");
  // one block column per matrix column, chunks of 256 rows
  for (int j0 = 0; j0 < n; j0 += 65535) {
    int nn = n - j0 < 65535 ? n - j0 : 65535;
    hipLaunchKernelGGL((copy2d_kernel<T>), dim3((m + 255) / 256, nn), dim3(256), 0, s, m, nn, X + (long)j0 * ldx,
                       ldx, Y + (long)j0 * ldy, ldy);
  }
  XK_HIP(hipGetLastError());
}

template <typename T>
__global__ void zero_tri_kernel(int uplo, int n, T* A, long lda)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n && j < n) {
    if ((uplo == Lower && i < j) || (uplo == Upper && i > j)) A[i + (long)j * lda] = T(0);
  }
}

template <typename T>
void zero_triangle(hipStream_t s, int uplo, int n, T* A, long lda)
{
  if (n <= 0) return;
  XK_CHECK(n <= 65535, "zero_triangle: order too large");
  hipLaunchKernelGGL((zero_tri_kernel<T>), dim3((n + 255) / 256, n), dim3(256), 0, s, uplo, n, A, lda);
  XK_HIP(hipGetLastError());
}

__global__ void fill_int_kernel(int* p, int n, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_int(hipStream_t s, int* p, int n, int v)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  XK_HIP(hipGetLastError());
}

#define XK_INST(T)                                                                                           \
  template void potrf_leaf<T>(hipStream_t, int, int, T*, long, long, T*, long, long, int, int*, int);         \
  template void trtri_leaf<T>(hipStream_t, int, int, int, const T*, long, long, T*, long, long, int);         \
  template void copy2d<T>(hipStream_t, int, int, const T*, long, T*, long);                                    \
  template void zero_triangle<T>(hipStream_t, int, int, T*, long);
XK_INST(double)
XK_INST(float)
#undef XK_INST

}  // namespace xk
