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

// Cholesky factor and its inverse for one block of order <= 64, blocked by
// 16 and spread over four waves.  The block and its inverse live in LDS
// (column-major, stride LLD).  For each 16-column block jb:
//   wave 0 factors the 16 x 16 diagonal block in registers (lane i holds row
//   i, cross-lane reads for the broadcasts: at most 15 per step) and forms
//   its inverse (lane c, column c), which is also the diagonal block of
//   inv(L);
//   all threads solve the panel below, L21 = A21 inv(L11)^T, and update the
//   trailing triangle, A22 -= L21 L21^T.
// The off-diagonal blocks of inv(L) follow block row by block row:
//   inv(L)[ib][jb] = -inv(L_ib,ib) sum_{t=jb}^{ib-1} L[ib][t] inv(L)[t][jb].
// Rows/columns beyond n are the identity.  info receives the first order
// whose pivot is not positive.
constexpr int LB = 16;  // inner block of the leaf

template <typename T>
__global__ __launch_bounds__(256) void potrf_leaf_kernel(int uplo, int n, T* A, long lda, long sA, T* Tinv,
                                                         long ldinv, long sInv, int* info, int base)
{
  __shared__ T Ls[LEAF * LLD];  // the factor
  __shared__ T Is[LEAF * LLD];  // its inverse
  __shared__ T Ts[3 * LB * LB]; // block-row partial products of the inverse
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  A += (long)blockIdx.x * sA;
  Tinv += (long)blockIdx.x * sInv;
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, k = e >> 6;
    Ls[i + k * LLD] = lower_view(A, lda, uplo, n, i, k);
    Is[i + k * LLD] = T(0);
  }
  __syncthreads();

  bool failed = false;  // (wave 0)
  for (int jb = 0; jb < LEAF / LB; ++jb) {
    const int c0 = jb * LB;
    if (wave == 0) {
      const int i = lane & (LB - 1);
      T d[LB];
#pragma unroll
      for (int k = 0; k < LB; ++k) d[k] = Ls[(c0 + i) + (c0 + k) * LLD];
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const T p = bcast(d[j], j);
        if (!(p > T(0)) && !failed && c0 + j < n) {
          failed = true;
          if (lane == 0) atomicMin(info, base + c0 + j + 1);
        }
        const T sq = sqrt(p);
        d[j] = i > j ? d[j] / sq : (i == j ? sq : T(0));
#pragma unroll
        for (int k = j + 1; k < LB; ++k) d[k] = fma(-d[j], bcast(d[j], k), d[k]);
      }
      T y[LB];  // column i of the block's inverse
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        T acc = (r == i) ? T(1) : T(0);
#pragma unroll
        for (int k = 0; k < r; ++k) acc = fma(-bcast(d[k], r), y[k], acc);
        y[r] = r < i ? T(0) : acc / bcast(d[r], r);
      }
      if (lane < LB) {
#pragma unroll
        for (int k = 0; k < LB; ++k) Ls[(c0 + lane) + (c0 + k) * LLD] = k <= lane ? d[k] : T(0);
#pragma unroll
        for (int r = 0; r < LB; ++r) Is[(c0 + r) + (c0 + lane) * LLD] = y[r];
      }
    }
    __syncthreads();
    const int m = LEAF - c0 - LB;  // rows below the diagonal block
    if (m > 0) {
      // panel: L[c0+LB+r][c0+c] = sum_{t <= c} A[c0+LB+r][c0+t] inv(L11)[c][t]
      T pv[3];
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int e = tid + 256 * u;
        pv[u] = T(0);
        if (e < m * LB) {
          const int r = e % m, c = e / m;
          // all 16 terms: inv(L11)[c][t] is 0 for t > c
          T acc = T(0);
#pragma unroll
          for (int t = 0; t < LB; ++t)
            acc = fma(Ls[(c0 + LB + r) + (c0 + t) * LLD], Is[(c0 + c) + (c0 + t) * LLD], acc);
          pv[u] = acc;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int e = tid + 256 * u;
        if (e < m * LB) Ls[(c0 + LB + e % m) + (c0 + e / m) * LLD] = pv[u];
      }
      __syncthreads();
      // trailing triangle: L[c0+LB+r][c0+LB+q] -= sum_t L21[r][t] L21[q][t], r >= q
      for (int e = tid; e < m * m; e += 256) {
        const int r = e % m, q = e / m;
        if (r >= q) {
          T acc = Ls[(c0 + LB + r) + (c0 + LB + q) * LLD];
#pragma unroll
          for (int t = 0; t < LB; ++t)
            acc = fma(-Ls[(c0 + LB + r) + (c0 + t) * LLD], Ls[(c0 + LB + q) + (c0 + t) * LLD], acc);
          Ls[(c0 + LB + r) + (c0 + LB + q) * LLD] = acc;
        }
      }
      __syncthreads();
    }
  }
  // factor back to the stored triangle
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, k = e >> 6;
    if (i < n && k < n && k <= i) {
      if (uplo == Lower) A[i + (long)k * lda] = Ls[i + k * LLD];
      else A[k + (long)i * lda] = Ls[i + k * LLD];
    }
  }
  if (!Tinv) return;
  for (int ib = 1; ib < LEAF / LB; ++ib) {
    // Ts[jb] = sum_{t=jb}^{ib-1} L[ib][t] inv(L)[t][jb], jb < ib
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int e = tid + 256 * u;
      if (e < ib * LB * LB) {
        const int jb = e >> 8, r = e & (LB - 1), c = (e >> 4) & (LB - 1);
        T acc = T(0);
        for (int tb = jb; tb < ib; ++tb) {
#pragma unroll
          for (int t = 0; t < LB; ++t) {
            const int s = tb * LB + t;
            acc = fma(Ls[(ib * LB + r) + s * LLD], Is[s + (jb * LB + c) * LLD], acc);
          }
        }
        Ts[e] = acc;
      }
    }
    __syncthreads();
    // inv(L)[ib][jb] = -inv(L_ib,ib) Ts[jb]
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int e = tid + 256 * u;
      if (e < ib * LB * LB) {
        const int jb = e >> 8, r = e & (LB - 1), c = (e >> 4) & (LB - 1);
        T acc = T(0);  // all 16 terms: the diagonal block's inverse is 0 above its diagonal
#pragma unroll
        for (int s = 0; s < LB; ++s)
          acc = fma(Is[(ib * LB + r) + (ib * LB + s) * LLD], Ts[(jb << 8) + s + (c << 4)], acc);
        Is[(ib * LB + r) + (jb * LB + c) * LLD] = -acc;
      }
    }
    __syncthreads();
  }
  // column j of inv(L) (zeros above the diagonal).  Lower: Tinv[i + j*ld];
  // Upper (inv(U) = inv(L)^T): Tinv[j + i*ld].
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, j = e >> 6;
    if (i < n && j < n) {
      if (uplo == Lower) Tinv[i + (long)j * ldinv] = Is[i + j * LLD];
      else Tinv[j + (long)i * ldinv] = Is[i + j * LLD];
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
  hipLaunchKernelGGL((potrf_leaf_kernel<T>), dim3(batch), dim3(256), 0, s, uplo, n, A, lda, sA, Tinv, ldinv, sInv,
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
