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
// 16 and spread over four waves.
// LDS: the lower triangle packed by rows, P[i (i + 1) / 2 + k] = L(i, k),
// and one 64 x 16 block column of the inverse: 24.3 KB in FP64.  A
// trailing-update GEMM that fills a CU (four 33 KB workgroups) leaves 28 KB
// of LDS free, so a leaf launched beside it starts at once in that gap
// instead of waiting for a CU to drain (a leaf holding the block and its
// inverse unpacked, 72 KB, waited up to milliseconds).
// For each 16-column block jb:
//   wave 0 factors the 16 x 16 diagonal block in registers (lane i holds row
//   i, cross-lane reads for the broadcasts: at most 15 per step), writes it
//   to A and puts its inverse (lane c, column c) in its place in P;
//   all threads solve the panel below, L21 = A21 inv(L11)^T, and update the
//   trailing triangle, A22 -= L21 L21^T.
// Then P holds L below the diagonal blocks and the diagonal blocks'
// inverses; inv(L) follows by block columns jb: X_jb = inv(L_jb,jb), X_ib =
// -inv(L_ib,ib) sum_{t=jb}^{ib-1} L[ib][t] X_t.  Rows/columns beyond n are the
// identity.  info receives the first order whose pivot is not positive.
constexpr int LB = 16;  // inner block of the leaf

__device__ __forceinline__ int pk(int i, int k) { return (i * (i + 1) >> 1) + k; }

template <typename T>
__global__ __launch_bounds__(256) void potrf_leaf_kernel(int uplo, int n, T* A, long lda, long sA, T* Tinv,
                                                         long ldinv, long sInv, int* info, int base, int pair)
{
  __shared__ T P[LEAF * (LEAF + 1) / 2];
  __shared__ T X[LEAF * LB];  // X[i * LB + c]: rows of one block column of inv(L)
  // the leaf is the latency-bound step of the diagonal chain and shares its
  // CU with update GEMM waves: its waves issue first
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  A += (long)blockIdx.x * sA;
  Tinv += (long)blockIdx.x * sInv;
  auto put = [&](int i, int k, T v) {  // factor entry (i, k), k <= i, to the stored triangle
    if (i < n && k < n) {
      if (uplo == Lower) A[i + (long)k * lda] = v;
      else A[k + (long)i * lda] = v;
    }
  };
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, k = e >> 6;
    if (k <= i) P[pk(i, k)] = lower_view(A, lda, uplo, n, i, k);
  }
  __syncthreads();

  bool failed = false;  // (wave 0)
  for (int jb = 0; jb < LEAF / LB; ++jb) {
    const int c0 = jb * LB;
    if (wave == 0) {
      const int i = lane & (LB - 1);
      T d[LB];
#pragma unroll
      for (int k = 0; k < LB; ++k) d[k] = k <= i ? P[pk(c0 + i, c0 + k)] : T(0);
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const T p = bcast(d[j], j);
        if (!(p > T(0)) && !failed && c0 + j < n) {
          failed = true;
          // pair: rows 2i, 2i+1 stand for order i of a complex block
          if (lane == 0) atomicMin(info, pair ? (base + c0 + j) / 2 + 1 : base + c0 + j + 1);
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
        for (int k = 0; k < LB; ++k)
          if (k <= lane) put(c0 + lane, c0 + k, d[k]);
        // (the whole block was read into d before any lane writes)
#pragma unroll
        for (int r = 0; r < LB; ++r)
          if (r >= lane) P[pk(c0 + r, c0 + lane)] = y[r];
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
          const T* arow = P + pk(c0 + LB + r, c0);
          const T* irow = P + pk(c0 + c, c0);
          T acc = T(0);
          for (int t = 0; t <= c; ++t) acc = fma(arow[t], irow[t], acc);
          pv[u] = acc;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int e = tid + 256 * u;
        if (e < m * LB) P[pk(c0 + LB + e % m, c0 + e / m)] = pv[u];
      }
      __syncthreads();
      // trailing triangle: L[c0+LB+r][c0+LB+q] -= sum_t L21[r][t] L21[q][t], r >= q
      for (int e = tid; e < m * m; e += 256) {
        const int r = e % m, q = e / m;
        if (r >= q) {
          const T* rr = P + pk(c0 + LB + r, c0);
          const T* rq = P + pk(c0 + LB + q, c0);
          T acc = P[pk(c0 + LB + r, c0 + LB + q)];
#pragma unroll
          for (int t = 0; t < LB; ++t) acc = fma(-rr[t], rq[t], acc);
          P[pk(c0 + LB + r, c0 + LB + q)] = acc;
        }
      }
      __syncthreads();
    }
  }
  // the factor below the diagonal blocks to the stored triangle
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, k = e >> 6;
    if (k < (i & ~(LB - 1))) put(i, k, P[pk(i, k)]);
  }
  if (!Tinv) return;
  // inv(L) by block columns; thread (r, c) owns entry (r, c) of a 16 x 16 block
  const int r = tid & (LB - 1), c = tid >> 4;
  auto out = [&](int i, int j, T v) {  // inv(L)(i, j) to Tinv (Upper: inv(U) = inv(L)^T)
    if (i < n && j < n) {
      if (uplo == Lower) Tinv[i + (long)j * ldinv] = v;
      else Tinv[j + (long)i * ldinv] = v;
    }
  };
  for (int jb = 0; jb < LEAF / LB; ++jb) {
    const int j0 = jb * LB;
    const T xd = c <= r ? P[pk(j0 + r, j0 + c)] : T(0);  // inv(L_jb,jb)
    X[(j0 + r) * LB + c] = xd;
    out(j0 + r, j0 + c, xd);
    for (int ib = 0; ib < jb; ++ib) out(ib * LB + r, j0 + c, T(0));
    __syncthreads();
    for (int ib = jb + 1; ib < LEAF / LB; ++ib) {
      const int i0 = ib * LB;
      T acc = T(0);
      const T* lrow = P + pk(i0 + r, 0);
      for (int t = j0; t < i0; ++t) acc = fma(lrow[t], X[t * LB + c], acc);
      X[(i0 + r) * LB + c] = acc;
      __syncthreads();
      T v = T(0);
      const T* drow = P + pk(i0 + r, i0);
      for (int s2 = 0; s2 <= r; ++s2) v = fma(drow[s2], X[(i0 + s2) * LB + c], v);
      __syncthreads();
      X[(i0 + r) * LB + c] = -v;
      out(i0 + r, j0 + c, -v);
      __syncthreads();
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
                int batch, int* info, int base, int pair)
{
  XK_CHECK(n >= 1 && n <= LEAF, "potrf leaf order must be in [1,64]");
  if (batch <= 0) return;
  hipLaunchKernelGGL((potrf_leaf_kernel<T>), dim3(batch), dim3(256), 0, s, uplo, n, A, lda, sA, Tinv, ldinv, sInv,
                     info, base, pair);
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
  template void potrf_leaf<T>(hipStream_t, int, int, T*, long, long, T*, long, long, int, int*, int, int);         \
  template void trtri_leaf<T>(hipStream_t, int, int, int, const T*, long, long, T*, long, long, int);         \
  template void copy2d<T>(hipStream_t, int, int, const T*, long, T*, long);                                    \
  template void zero_triangle<T>(hipStream_t, int, int, T*, long);
XK_INST(double)
XK_INST(float)
#undef XK_INST

}  // namespace xk
