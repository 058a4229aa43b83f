// complex.hip — complex GEMM on the real MFMA kernel.
//
// With X = op(A) (m x k) and Y = op(B) (k x n) complex, the product P = X Y
// is one real GEMM of twice the order:
//     [Pr]   [Xr  -Xi] [Yr]
//     [Pi] = [Xi   Xr] [Yi]      (2m x 2k) (2k x n) -> (2m x n)
// (the "4M" form: four real products, here in a single launch).  These
// kernels build the real operands from interleaved (re, im) storage — for a
// transposed or conjugated operand they build the transpose of the block
// matrix, which the GEMM then reads transposed — and fold alpha, beta back
// into the interleaved C.  All are single HBM passes.
#include "kernels.h"
#include <algorithm>

namespace xk {

template <typename T>
struct Cplx {
  T re, im;
};

// Y (2 rows x 2 cols) <- [Ar, -s Ai; s Ai, Ar] for the rows x cols complex A.
template <typename T>
__global__ __launch_bounds__(256) void cplx_pack_a_kernel(int rows, int cols, int j0, T s, const Cplx<T>* __restrict__ A,
                                                          long lda, T* __restrict__ Y, long ldy)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= rows || j >= cols) return;
  const Cplx<T> a = A[i + (long)j * lda];
  Y[i + (long)j * ldy] = a.re;
  Y[i + rows + (long)j * ldy] = s * a.im;
  Y[i + (long)(j + cols) * ldy] = -s * a.im;
  Y[i + rows + (long)(j + cols) * ldy] = a.re;
}

// vertical: Y (2 rows x cols) <- [Br; Bi]; else Y (rows x 2 cols) <- [Br, s Bi].
template <typename T>
__global__ __launch_bounds__(256) void cplx_pack_b_kernel(int vertical, int rows, int cols, int j0, T s,
                                                          const Cplx<T>* __restrict__ B, long ldb, T* __restrict__ Y,
                                                          long ldy)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= rows || j >= cols) return;
  const Cplx<T> b = B[i + (long)j * ldb];
  Y[i + (long)j * ldy] = b.re;
  if (vertical) Y[i + rows + (long)j * ldy] = b.im;
  else Y[i + (long)(j + cols) * ldy] = s * b.im;
}

// Two real blocks of the rows x cols complex B side by side (vertical = 0:
// Y rows x 2 cols) or stacked (vertical: Y 2 rows x cols): [sr Br, si Bi],
// or with swap [si Bi, sr Br].
template <typename T>
__global__ __launch_bounds__(256) void cplx_pack2_kernel(int vertical, int swap, T sr, T si, int rows, int cols,
                                                         int j0, const Cplx<T>* __restrict__ B, long ldb,
                                                         T* __restrict__ Y, long ldy)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= rows || j >= cols) return;
  const Cplx<T> b = B[i + (long)j * ldb];
  const T first = swap ? si * b.im : sr * b.re, second = swap ? sr * b.re : si * b.im;
  Y[i + (long)j * ldy] = first;
  if (vertical) Y[i + rows + (long)j * ldy] = second;
  else Y[i + (long)(j + cols) * ldy] = second;
}

// C <- alpha (Pr + i Pi) + beta C, P = [Pr; Pi] real (2m x n); use_p = 0:
// C <- beta C; beta = 0 never reads C.  uplo Lower / Upper: only that
// triangle of the square C, with a real diagonal when herm (HERK).
template <typename T>
__global__ __launch_bounds__(256) void cplx_combine_kernel(int m, int n, int j0, int use_p, const T* __restrict__ P,
                                                           long ldp, T ar, T ai, T br, T bi, Cplx<T>* __restrict__ C,
                                                           long ldc, int uplo, int herm)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= m || j >= n) return;
  if ((uplo == Lower && i < j) || (uplo == Upper && i > j)) return;
  T re = T(0), im = T(0);
  if (use_p) {
    const T pr = P[i + (long)j * ldp], pi = P[i + m + (long)j * ldp];
    re = ar * pr - ai * pi;
    im = ar * pi + ai * pr;
  }
  Cplx<T>* c = C + i + (long)j * ldc;
  if (br != T(0) || bi != T(0)) {
    const Cplx<T> v = *c;
    re += br * v.re - bi * v.im;
    im += br * v.im + bi * v.re;
  }
  if (herm && i == j) im = T(0);
  *c = Cplx<T>{re, im};
}

// Y (n x n) <- the matrix whose uplo triangle is stored in A: mode 0
// symmetric, 1 Hermitian (mirror conjugated, imaginary part of the diagonal
// taken as zero), 2 triangular (zeros outside), 3 unit triangular (diagonal
// of ones, not read).
template <typename T>
__global__ __launch_bounds__(256) void cplx_expand_kernel(int mode, int uplo, int n, int j0,
                                                          const Cplx<T>* __restrict__ A, long lda,
                                                          Cplx<T>* __restrict__ Y, long ldy)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= n || j >= n) return;
  const bool stored = (uplo == Lower) == (i > j);
  Cplx<T> v;
  if (i == j) {
    if (mode == 3) v = Cplx<T>{T(1), T(0)};
    else {
      v = A[i + (long)i * lda];
      if (mode == 1) v.im = T(0);
    }
  } else if (stored) {
    v = A[i + (long)j * lda];
  } else if (mode >= 2) {
    v = Cplx<T>{T(0), T(0)};
  } else {
    v = A[j + (long)i * lda];
    if (mode == 1) v.im = -v.im;
  }
  Y[i + (long)j * ldy] = v;
}

// Inverse of `batch` complex triangular blocks of order n <= 64 (block b at
// A + b sA, inverse at Tinv + b sInv, complex elements).  One wave per block,
// lane j solving column j of inv(L) by forward substitution in the lower view
// (Upper: L = U^T, inv(U) = inv(L)^T); L is read through the cache (all lanes
// read the same entry), the columns being built live in LDS.
template <typename T>
__global__ __launch_bounds__(64) void cplx_trtri_leaf_kernel(int uplo, int unit, int n, const Cplx<T>* __restrict__ A,
                                                             long lda, long sA, Cplx<T>* __restrict__ Tinv, long ldinv,
                                                             long sInv)
{
  __shared__ T Xr[64 * 65], Xi[64 * 65];
  const int j = threadIdx.x;
  A += (long)blockIdx.x * sA;
  Tinv += (long)blockIdx.x * sInv;
  auto L = [&](int i, int k) { return uplo == Lower ? A[i + (long)k * lda] : A[k + (long)i * lda]; };
  for (int i = 0; i < n; ++i) {
    // s = sum_{k < i} L(i,k) x_k; x_k = 0 for k < j, so all lanes run the same loop
    T sr = T(0), si = T(0);
    for (int k = 0; k < i; ++k) {
      const Cplx<T> l = L(i, k);
      const T xr = Xr[k * 65 + j], xi = Xi[k * 65 + j];
      sr += l.re * xr - l.im * xi;
      si += l.re * xi + l.im * xr;
    }
    T vr, vi;
    if (i < j) {
      vr = T(0);
      vi = T(0);
    } else {
      // i == j: x = 1 / L(i,i); i > j: x = -s / L(i,i)
      const T nr = i == j ? T(1) : -sr, ni = i == j ? T(0) : -si;
      if (unit) {
        vr = nr;
        vi = ni;
      } else {
        const Cplx<T> d = L(i, i);
        const T den = d.re * d.re + d.im * d.im;
        vr = (nr * d.re + ni * d.im) / den;
        vi = (ni * d.re - nr * d.im) / den;
      }
    }
    if (j < n) {
      Xr[i * 65 + j] = vr;
      Xi[i * 65 + j] = vi;
    }
    __syncthreads();
  }
  if (j >= n) return;
  for (int i = 0; i < n; ++i) {
    const Cplx<T> v{Xr[i * 65 + j], Xi[i * 65 + j]};
    if (uplo == Lower) Tinv[i + (long)j * ldinv] = v;
    else Tinv[j + (long)i * ldinv] = v;
  }
}

// Real form of a Hermitian block for its Cholesky factorization: the lower
// triangle (r >= c) of the 2n x 2n real M with M[2i+p][2j+q] = rho(a_ij)[p][q],
// rho(z) = [re -im; im re], a_ij (i >= j) from the stored triangle (Upper:
// conj(A(j, i))); the diagonal's imaginary parts are taken as zero, as LAPACK
// does.
template <typename T>
__global__ __launch_bounds__(256) void cplx_realform_kernel(int uplo, int n, int c0, const Cplx<T>* __restrict__ A,
                                                            long lda, T* __restrict__ R, long ldr)
{
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c = c0 + (int)blockIdx.y;
  if (r >= 2 * n || c > r) return;
  const int i = r >> 1, j = c >> 1, p = r & 1, q = c & 1;
  Cplx<T> a = uplo == Lower ? A[i + (long)j * lda] : A[j + (long)i * lda];
  if (uplo != Lower) a.im = -a.im;
  if (i == j) a.im = T(0);
  R[r + (long)c * ldr] = p == q ? a.re : (p ? a.im : -a.im);
}

// The complex factor and inverse back from the real ones: l_ij = R[2i][2j] +
// i R[2i+1][2j] for i >= j (a real diagonal), to A's stored triangle (Upper:
// U(j, i) = conj(l_ij)); likewise inv(L) from Rinv to Tinv (Upper: inv(U) =
// inv(L)^H).  Tinv's other triangle is not written.
template <typename T>
__global__ __launch_bounds__(256) void cplx_fromreal_kernel(int uplo, int n, int j0, const T* __restrict__ R, long ldr,
                                                            Cplx<T>* __restrict__ A, long lda,
                                                            const T* __restrict__ Rinv, Cplx<T>* __restrict__ Tinv,
                                                            long ldinv)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = j0 + (int)blockIdx.y;
  if (i >= n || j > i) return;
  const long e = 2 * i + (long)(2 * j) * ldr;
  Cplx<T> l{R[e], i == j ? T(0) : R[e + 1]};
  Cplx<T> v{Rinv[e], i == j ? T(0) : Rinv[e + 1]};
  if (uplo == Lower) {
    A[i + (long)j * lda] = l;
    Tinv[i + (long)j * ldinv] = v;
  } else {
    A[j + (long)i * lda] = Cplx<T>{l.re, -l.im};
    Tinv[j + (long)i * ldinv] = Cplx<T>{v.re, -v.im};
  }
}

template <typename F>
static void over_columns(int rows, int cols, F launch)
{
  for (int j0 = 0; j0 < cols; j0 += 65535) launch(dim3((rows + 255) / 256, std::min(65535, cols - j0)), j0);
}

template <typename T>
void cplx_pack_a(hipStream_t s, int sigma, int rows, int cols, const T* A, long lda, T* Y, long ldy)
{
  if (rows <= 0 || cols <= 0) return;
  over_columns(rows, cols, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_pack_a_kernel<T>), g, dim3(256), 0, s, rows, cols, j0, T(sigma),
                       (const Cplx<T>*)A, lda, Y, ldy);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_pack_b(hipStream_t s, int vertical, int sign, int rows, int cols, const T* B, long ldb, T* Y, long ldy)
{
  if (rows <= 0 || cols <= 0) return;
  over_columns(rows, cols, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_pack_b_kernel<T>), g, dim3(256), 0, s, vertical, rows, cols, j0, T(sign),
                       (const Cplx<T>*)B, ldb, Y, ldy);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_pack2(hipStream_t s, int vertical, int swap, int sr, int si, int rows, int cols, const T* B, long ldb, T* Y,
                long ldy)
{
  if (rows <= 0 || cols <= 0) return;
  over_columns(rows, cols, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_pack2_kernel<T>), g, dim3(256), 0, s, vertical, swap, T(sr), T(si), rows, cols, j0,
                       (const Cplx<T>*)B, ldb, Y, ldy);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_combine(hipStream_t s, int m, int n, const T* P, long ldp, const T* alpha, const T* beta, T* C, long ldc,
                  int uplo, int herm)
{
  if (m <= 0 || n <= 0) return;
  over_columns(m, n, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_combine_kernel<T>), g, dim3(256), 0, s, m, n, j0, P != nullptr ? 1 : 0, P, ldp,
                       alpha[0], alpha[1], beta[0], beta[1], (Cplx<T>*)C, ldc, uplo, herm);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_expand(hipStream_t s, int mode, int uplo, int n, const T* A, long lda, T* Y, long ldy)
{
  if (n <= 0) return;
  over_columns(n, n, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_expand_kernel<T>), g, dim3(256), 0, s, mode, uplo, n, j0, (const Cplx<T>*)A, lda,
                       (Cplx<T>*)Y, ldy);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_trtri_leaf(hipStream_t s, int uplo, int unit, int n, const T* A, long lda, long sA, T* Tinv, long ldinv,
                     long sInv, int batch)
{
  XK_CHECK(n >= 1 && n <= 64, "complex trtri leaf order must be in [1,64]");
  if (batch <= 0) return;
  hipLaunchKernelGGL((cplx_trtri_leaf_kernel<T>), dim3(batch), dim3(64), 0, s, uplo, unit, n, (const Cplx<T>*)A, lda,
                     sA, (Cplx<T>*)Tinv, ldinv, sInv);
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_realform(hipStream_t s, int uplo, int n, const T* A, long lda, T* R, long ldr)
{
  if (n <= 0) return;
  over_columns(2 * n, 2 * n, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_realform_kernel<T>), g, dim3(256), 0, s, uplo, n, j0, (const Cplx<T>*)A, lda, R, ldr);
  });
  XK_HIP(hipGetLastError());
}

template <typename T>
void cplx_fromreal(hipStream_t s, int uplo, int n, const T* R, long ldr, T* A, long lda, const T* Rinv, T* Tinv,
                   long ldinv)
{
  if (n <= 0) return;
  over_columns(n, n, [&](dim3 g, int j0) {
    hipLaunchKernelGGL((cplx_fromreal_kernel<T>), g, dim3(256), 0, s, uplo, n, j0, R, ldr, (Cplx<T>*)A, lda, Rinv,
                       (Cplx<T>*)Tinv, ldinv);
  });
  XK_HIP(hipGetLastError());
}

#define XK_INST(T)                                                                                         \
  template void cplx_realform<T>(hipStream_t, int, int, const T*, long, T*, long);                         \
  template void cplx_fromreal<T>(hipStream_t, int, int, const T*, long, T*, long, const T*, T*, long);      \
  template void cplx_trtri_leaf<T>(hipStream_t, int, int, int, const T*, long, long, T*, long, long, int); \
  template void cplx_pack_a<T>(hipStream_t, int, int, int, const T*, long, T*, long);                      \
  template void cplx_pack_b<T>(hipStream_t, int, int, int, int, const T*, long, T*, long);                 \
  template void cplx_pack2<T>(hipStream_t, int, int, int, int, int, int, const T*, long, T*, long);        \
  template void cplx_combine<T>(hipStream_t, int, int, const T*, long, const T*, const T*, T*, long, int, int);  \
  template void cplx_expand<T>(hipStream_t, int, int, int, const T*, long, T*, long);
XK_INST(double)
XK_INST(float)
#undef XK_INST

}  // namespace xk
