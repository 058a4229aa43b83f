// factor.hip — latency-bound diagonal-block kernels for gfx950:
// Cholesky + inverse of blocks of order <= 64, triangular inverse, and small
// helpers (2-D copy, triangle zeroing, int fill).
//
// The padded rows/columns of a leaf beyond n are the identity, so every
// block is handled as a 64 x 64 problem.  The Cholesky leaf (four waves)
// writes the factor and its inverse in one launch; the triangular-inverse
// leaf (one wave) holds the block in LDS (column stride 65) and solves one
// column of inv(L) per lane.
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

// sq = sqrt(p) and rs = 1 / sq for a finite pivot p > 0, without a division
// and without a branch: p = 2^(2h) p' with p' in [0.5, 2) (exact scaling by
// the exponent, denormal p included), Goldschmidt's iteration on p' from the
// hardware reciprocal square root, two Newton corrections of the root
// (correctly rounded) and one of the reciprocal (about 1 ulp), then
// sq = 2^h sqrt(p'), rs = 2^-h / sqrt(p').  For p <= 0, infinite or NaN the
// results are unspecified (the caller reports such a pivot from p itself).
__device__ __forceinline__ void sqrt_rcp(double p, double& sq, double& rs)
{
  const int h = __builtin_amdgcn_frexp_exp(p) >> 1;
  const double ps = __builtin_ldexp(p, -2 * h);
  const double r = __builtin_amdgcn_rsq(ps);
  double g = ps * r, hh = 0.5 * r;
  const double e = fma(-hh, g, 0.5);
  g = fma(g, e, g);
  hh = fma(hh, e, hh);
  g = fma(fma(-g, g, ps), hh, g);
  g = fma(fma(-g, g, ps), hh, g);
  double q = 2.0 * hh;
  q = fma(q, fma(-g, q, 1.0), q);
  sq = __builtin_ldexp(g, h);
  rs = __builtin_ldexp(q, -h);
}
__device__ __forceinline__ void sqrt_rcp(float p, float& sq, float& rs)
{
  sq = sqrtf(p);
  rs = 1.0f / sq;
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

// Workgroup barrier that orders LDS accesses only: global stores in flight
// are not waited for (nothing in the kernel reads them back).
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Cholesky factor and its inverse for one block of order <= 64, blocked by
// 16 and spread over four waves.
// LDS: the lower triangle packed by rows, P[i (i + 1) / 2 + k] = L(i, k),
// and one 64 x 16 block column of the inverse: 24.4 KB in FP64.
// A trailing-update GEMM that fills a CU (four 33 KB workgroups) leaves
// 28 KB of LDS free, so a leaf launched beside it starts at once in that gap
// instead of waiting for a CU to drain (a leaf holding the block and its
// inverse unpacked, 72 KB, waited up to milliseconds).  About 125 VGPRs: it
// also fits beside three update-GEMM waves (104-112 VGPRs each) on a SIMD.
// For each 16-column block jb:
//   wave 0 factors the 16 x 16 diagonal block in registers (lane i holds row
//   i, cross-lane reads for the broadcasts), with the pivots' roots and
//   reciprocals from one reciprocal-square-root iteration (no division on
//   the chain), then inverts it (lane c solves column c), writes L to A and
//   the inverse in its place in P;
//   all threads solve the panel below, L21 = A21 inv(L11)^T, and update the
//   trailing triangle, A22 -= L21 L21^T.
// Then P holds L below the diagonal blocks and the diagonal blocks'
// inverses, from which inv(L) follows by block columns.  The barriers order
// LDS only (no wait for the global stores of L and inv(L)).  Rows/columns
// beyond n are the identity.  info receives the first order whose pivot is
// not positive.
constexpr int LB = 16;  // inner block of the leaf

__device__ __forceinline__ int pk(int i, int k) { return (i * (i + 1) >> 1) + k; }

template <typename T>
__global__ __launch_bounds__(256) void potrf_leaf_kernel(int uplo, int n, T* A, long lda, long sA, T* Tinv,
                                                         long ldinv, long sInv, int* info, int base, int pair)
{
  __shared__ T P[LEAF * (LEAF + 1) / 2];
  __shared__ T X[LEAF * LB];            // X[i * LB + c]: rows of one block column of inv(L)
  __shared__ T Lr[LB];                  // reciprocals of the diagonal block's pivots
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
  {
    // all loads in flight before the first LDS store
    T v[LEAF * LEAF / 256];
#pragma unroll
    for (int u = 0; u < LEAF * LEAF / 256; ++u) {
      const int e = tid + 256 * u, i = e & 63, k = e >> 6;
      v[u] = k <= i ? lower_view(A, lda, uplo, n, i, k) : T(0);
    }
#pragma unroll
    for (int u = 0; u < LEAF * LEAF / 256; ++u) {
      const int e = tid + 256 * u, i = e & 63, k = e >> 6;
      if (k <= i) P[pk(i, k)] = v[u];
    }
  }
  lds_barrier();

  int fail = INT_MAX;  // first order with a non-positive pivot (wave 0)
  for (int jb = 0; jb < LEAF / LB; ++jb) {
    const int c0 = jb * LB;
    if (wave == 0) {
      const int i = lane & (LB - 1);
      const bool own = lane < LB;
      T d[LB];
#pragma unroll
      for (int k = 0; k < LB; ++k) d[k] = k <= i ? P[pk(c0 + i, c0 + k)] : T(0);
      // right-looking: the pivot chain needs no cross-lane data beyond the
      // pivot itself (lane j + 1 updates its own diagonal entry).  The chain
      // has no branch (lane j keeps reciprocal j, stored after the loop), so
      // the updates of one step are scheduled under the next pivot's root.
      T myrs = T(0);
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const T p = bcast(d[j], j);
        fail = (!(p > T(0)) && fail == INT_MAX && c0 + j < n) ? c0 + j : fail;
        T sq, rs;
        sqrt_rcp(p, sq, rs);
        myrs = lane == j ? rs : myrs;
        d[j] = i > j ? d[j] * rs : (i == j ? sq : T(0));
#pragma unroll
        for (int k = j + 1; k < LB; ++k) d[k] = fma(-d[j], bcast(d[j], k), d[k]);
      }
      if (lane < LB) Lr[lane] = myrs;
      if (own) {
#pragma unroll
        for (int k = 0; k < LB; ++k)
          if (k <= i) put(c0 + i, c0 + k, d[k]);
      }
      // column i of inv(L11): y_r = (delta_ri - sum_{k<r} L(r, k) y_k) / L(r, r),
      // L(r, k) read from lane r
      T y[LB];
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        T acc = (r == i) ? T(1) : T(0);
#pragma unroll
        for (int k = 0; k < r; ++k) acc = fma(-bcast(d[k], r), y[k], acc);
        y[r] = r < i ? T(0) : acc * Lr[r];
      }
#pragma unroll
      for (int r = 0; r < LB; ++r) d[r] = y[r];
      if (own) {
#pragma unroll
        for (int r = 0; r < LB; ++r)
          if (r >= i) P[pk(c0 + r, c0 + i)] = d[r];
      }
    }
    lds_barrier();
    const int m = LEAF - c0 - LB;  // rows below the diagonal block
    if (m > 0) {
      // panel: L[c0+LB+r][c0+c] = sum_{t <= c} A[c0+LB+r][c0+t] inv(L11)[c][t]
      T pv[3];
#pragma unroll 1
      for (int u = 0; u < 3; ++u) {
        const int e = tid + 256 * u;
        pv[u] = T(0);
        if (e < m * LB) {
          const int r = e % m, c = e / m;
          const T* arow = P + pk(c0 + LB + r, c0);
          const T* irow = P + pk(c0 + c, c0);
          T a0 = T(0), a1 = T(0);
#pragma unroll
          for (int t = 0; t < LB; t += 2) {
            a0 = fma(arow[t], t <= c ? irow[t] : T(0), a0);
            a1 = fma(arow[t + 1], t + 1 <= c ? irow[t + 1] : T(0), a1);
          }
          pv[u] = a0 + a1;
        }
      }
      lds_barrier();
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int e = tid + 256 * u;
        if (e < m * LB) P[pk(c0 + LB + e % m, c0 + e / m)] = pv[u];
      }
      lds_barrier();
      // trailing triangle: L[c0+LB+r][c0+LB+q] -= sum_t L21[r][t] L21[q][t],
      // q <= r < m, element e = r (r + 1) / 2 + q of the triangle; each
      // element is read and written by its own thread only
      const int cnt = m * (m + 1) / 2;
#pragma unroll 1
      for (int u = 0; u < 5; ++u) {  // cnt <= 48 * 49 / 2 < 5 * 256
        const int e = tid + 256 * u;
        if (e < cnt) {
          int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
          if ((r + 1) * (r + 2) / 2 <= e) ++r;
          if (r * (r + 1) / 2 > e) --r;
          const int q = e - r * (r + 1) / 2;
          const T* rr = P + pk(c0 + LB + r, c0);
          const T* rq = P + pk(c0 + LB + q, c0);
          T a0 = P[pk(c0 + LB + r, c0 + LB + q)], a1 = T(0);
#pragma unroll
          for (int t = 0; t < LB; t += 2) {
            a0 = fma(-rr[t], rq[t], a0);
            a1 = fma(-rr[t + 1], rq[t + 1], a1);
          }
          P[pk(c0 + LB + r, c0 + LB + q)] = a0 + a1;
        }
      }
      lds_barrier();
    }
  }
  // pair: orders 2i, 2i + 1 stand for order i of a complex block
  if (tid == 0 && fail != INT_MAX) atomicMin(info, pair ? (base + fail) / 2 + 1 : base + fail + 1);
  // the factor below the diagonal blocks to the stored triangle
  for (int e = tid; e < LEAF * LEAF; e += 256) {
    const int i = e & 63, k = e >> 6;
    if (k < (i & ~(LB - 1))) put(i, k, P[pk(i, k)]);
  }
  if (!Tinv) return;
  auto out = [&](int i, int j, T v) {  // inv(L)(i, j) to Tinv (Upper: inv(U) = inv(L)^T)
    if (i < n && j < n) {
      if (uplo == Lower) Tinv[i + (long)j * ldinv] = v;
      else Tinv[j + (long)i * ldinv] = v;
    }
  };
  // inv(L) by block columns jb: X_jb = inv(L_jb,jb), X_ib = -inv(L_ib,ib)
  // sum_{t=jb}^{ib-1} L[ib][t] X_t; thread (r, c) owns entry (r, c) of a
  // 16 x 16 block
  const int r = tid & (LB - 1), c = tid >> 4;
  for (int jb = 0; jb < LEAF / LB; ++jb) {
    const int j0 = jb * LB;
    const T xd = c <= r ? P[pk(j0 + r, j0 + c)] : T(0);  // inv(L_jb,jb)
    X[(j0 + r) * LB + c] = xd;
    out(j0 + r, j0 + c, xd);
    for (int ib = 0; ib < jb; ++ib) out(ib * LB + r, j0 + c, T(0));
    lds_barrier();
    for (int ib = jb + 1; ib < LEAF / LB; ++ib) {
      const int i0 = ib * LB;
      const T* lrow = P + pk(i0 + r, 0);
      T a0 = T(0), a1 = T(0);
      for (int t = j0; t < i0; t += 2) {  // i0 - j0 is a multiple of 16
        a0 = fma(lrow[t], X[t * LB + c], a0);
        a1 = fma(lrow[t + 1], X[(t + 1) * LB + c], a1);
      }
      X[(i0 + r) * LB + c] = a0 + a1;
      lds_barrier();
      const T* drow = P + pk(i0 + r, i0);
      a0 = T(0);
      a1 = T(0);
#pragma unroll
      for (int s2 = 0; s2 < LB; s2 += 2) {
        a0 = fma(s2 <= r ? drow[s2] : T(0), X[(i0 + s2) * LB + c], a0);
        a1 = fma(s2 + 1 <= r ? drow[s2 + 1] : T(0), X[(i0 + s2 + 1) * LB + c], a1);
      }
      lds_barrier();
      X[(i0 + r) * LB + c] = -(a0 + a1);
      out(i0 + r, j0 + c, -(a0 + a1));
      lds_barrier();
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

// C <- beta C + sum_p W_p (W_p: m x n, leading dimension m, p-th at p m n);
// beta == 0 never reads C.
template <typename T>
__global__ void splitk_reduce_kernel(int m, int n, int parts, const T* __restrict__ W, T beta, T* __restrict__ C,
                                     long ldc)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  const long mn = (long)m * n, e = i + (long)j * m;
  T v = W[e];
  for (int p = 1; p < parts; ++p) v += W[p * mn + e];
  T* c = C + i + (long)j * ldc;
  *c = beta != T(0) ? beta * *c + v : v;
}

template <typename T>
void splitk_reduce(hipStream_t s, int m, int n, int parts, const T* W, T beta, T* C, long ldc)
{
  if (m <= 0 || n <= 0) return;
  XK_CHECK(n <= 65535, "splitk_reduce: too many columns");
  hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, parts, W, beta, C,
                     ldc);
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
  template void zero_triangle<T>(hipStream_t, int, int, T*, long);                                            \
  template void splitk_reduce<T>(hipStream_t, int, int, int, const T*, T, T*, long);
XK_INST(double)
XK_INST(float)
#undef XK_INST

}  // namespace xk
