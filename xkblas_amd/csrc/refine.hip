// refine.hip — element-wise kernels around the GEMM: expansion of one stored
// triangle to the full symmetric matrix (SYMM, the SPD solve's residual) or to
// a full triangular one (TRMM), column max-norms and absolute sums, X += D.
// All are HBM-bound single passes with coalesced column-major accesses.
#include "kernels.h"
#include <algorithm>

namespace xk {

// Y (and Z when non-null) <- the symmetric matrix whose uplo triangle is
// stored in A.  One 64 x 64 tile per workgroup (256 threads as 64 rows x 4
// column phases); a tile on the mirrored side reads its transposed source
// tile coalesced and transposes it through LDS.
template <typename T>
__global__ __launch_bounds__(256) void symm_expand_kernel(int uplo, int n, const T* __restrict__ A, long lda,
                                                          T* __restrict__ Y, long ldy, T* __restrict__ Z, long ldz)
{
  __shared__ T t[64][65];
  const int bi = blockIdx.x, bj = blockIdx.y;
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

template <typename T>
void symm_expand(hipStream_t s, int uplo, int n, const T* A, long lda, T* Y, long ldy, T* Z, long ldz)
{
  if (n <= 0) return;
  const int nt = (n + 63) / 64;
  XK_CHECK(nt < 65536, "symm_expand: order too large");
  hipLaunchKernelGGL((symm_expand_kernel<T>), dim3(nt, nt), dim3(256), 0, s, uplo, n, A, lda, Y, ldy, Z, ldz);
  XK_HIP(hipGetLastError());
}
template void symm_expand<double>(hipStream_t, int, int, const double*, long, double*, long, double*, long);
template void symm_expand<float>(hipStream_t, int, int, const float*, long, float*, long, float*, long);

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
