// kernels.h — host-side launchers of the hand-written gfx950 kernels.
// All pointers are device pointers into column-major buffers; every launch is
// asynchronous on the given stream.
#pragma once
#include "common.h"

namespace xk {

// C <- alpha op(A) op(B) + beta C, op(A) m x k, op(B) k x n.  `tri` restricts
// the written part of C to one triangle (C's (0,0) lies on the diagonal).
// `batch` > 1 runs independent problems at element strides sA, sB, sC.
// beta == 0 never reads C.  FP64 uses v_mfma_f64_16x16x4f64, FP32 uses
// v_mfma_f32_16x16x4f32; tiles are staged through LDS, double-buffered.
// `kmode` (KRANGE_*) limits each tile's inner range to where a triangular
// op(A) (its rows) or op(B) (its columns) is not zero.
template <typename T>
void gemm(hipStream_t s, int transA, int transB, int m, int n, int k, T alpha,
          const T* A, long lda, const T* B, long ldb, T beta, T* C, long ldc,
          int tri = TRI_FULL, int batch = 1, long sA = 0, long sB = 0, long sC = 0, int kmode = 0);

// Forces the macro-tile shape of gemm(): 0 = automatic, 32 (FP64; FP32 runs
// 64), 64 or 128.
void gemm_set_tile(int bm);
int gemm_get_tile();

// Cholesky of `batch` diagonal blocks of order n <= 64 (lower storage, or
// upper when uplo == Upper), in place, plus the inverse of each factor
// written to Linv (lower: inv(L); upper: inv(U)).  info[0] receives
// base_order + j + 1 for the first non-positive pivot (atomic min, starts at
// INT_MAX), nothing on success.  pair != 0: the block is the real form of a
// complex block (orders 2i, 2i + 1 stand for complex order i), and a failing
// pivot j reports (base_order + j) / 2 + 1.
template <typename T>
void potrf_leaf(hipStream_t s, int uplo, int n, T* A, long lda, long strideA, T* Linv, long ldinv,
                long strideInv, int batch, int* info, int base_order, int pair = 0);

// Inverse of `batch` triangular blocks of order n <= 64 into Tinv (the other
// triangle of Tinv is zeroed, like the full block).
template <typename T>
void trtri_leaf(hipStream_t s, int uplo, int diag, int n, const T* A, long lda, long strideA, T* Tinv,
                long ldinv, long strideInv, int batch);

// Y <- X (m x n blocks), element-wise copy between column-major buffers.
template <typename T>
void copy2d(hipStream_t s, int m, int n, const T* X, long ldx, T* Y, long ldy);

// Zero the strictly upper (uplo == Lower) or strictly lower part of an n x n
// block.
template <typename T>
void zero_triangle(hipStream_t s, int uplo, int n, T* A, long lda);

// C <- beta C + sum over `parts` of the m x n partial products W (leading
// dimension m, one every m n elements); beta == 0 never reads C.
template <typename T>
void splitk_reduce(hipStream_t s, int m, int n, int parts, const T* W, T beta, T* C, long ldc);

// Fill an int array with a value (stream-ordered, graph-capturable).
void fill_int(hipStream_t s, int* p, int n, int v);
// Products enqueued on stream s (null: none) raise their waves' issue
// priority: the latency-bound diagonal chain of POTRF beside full-chip
// trailing updates.
void gemm_set_priority_stream(hipStream_t s);
// device address of the allocation-free batched run's fixed array, for the
// current device (called once per device at initialisation)
void batch_noalloc_init();

// Y (and Z when non-null) <- the full symmetric n x n matrix whose uplo
// triangle is stored in A, columns c0 .. c1 - 1 (c0 a multiple of 64; c1 < 0:
// up to n; whole tiles of 64 columns are written).
template <typename T>
void symm_expand(hipStream_t s, int uplo, int n, const T* A, long lda, T* Y, long ldy, T* Z, long ldz, int c0 = 0,
                 int c1 = -1);

// Y <- the n x n triangular matrix whose uplo triangle is stored in A (zeros
// outside it, a unit diagonal when unit != 0).
template <typename T>
void tri_expand(hipStream_t s, int uplo, int unit, int n, const T* A, long lda, T* Y, long ldy);

// out[j] <- max_i |X(i, j)| (op 0, NaN-propagating) or sum_i |X(i, j)| (op 1)
// for the n columns of the m x n block X.
void col_reduce(hipStream_t s, int op, int m, int n, const double* X, long ldx, double* out);

// Complex GEMM operands for one real GEMM of twice the order (complex
// matrices interleaved (re, im), leading dimensions in complex elements):
// Y (2 rows x 2 cols) <- [Ar, -sigma Ai; sigma Ai, Ar];
// Y <- [Br; Bi] (vertical, 2 rows x cols) or [Br, sign Bi] (rows x 2 cols);
// C <- alpha (Pr + i Pi) + beta C with P = [Pr; Pi] (2m x n; P null: C <-
// beta C; beta 0 never reads C); alpha, beta as (re, im).
template <typename T>
void cplx_pack_a(hipStream_t s, int sigma, int rows, int cols, const T* A, long lda, T* Y, long ldy);
template <typename T>
void cplx_pack_b(hipStream_t s, int vertical, int sign, int rows, int cols, const T* B, long ldb, T* Y, long ldy);
// Y <- [sr Br, si Bi] (rows x 2 cols) or, vertical, [sr Br; si Bi] (2 rows x
// cols); swap puts the si Bi block first.
template <typename T>
void cplx_pack2(hipStream_t s, int vertical, int swap, int sr, int si, int rows, int cols, const T* B, long ldb, T* Y,
                long ldy);
// uplo Lower / Upper: only that triangle of the square C is written, with a
// real diagonal when herm (HERK); 0: all of C.
template <typename T>
void cplx_combine(hipStream_t s, int m, int n, const T* P, long ldp, const T* alpha, const T* beta, T* C, long ldc,
                  int uplo = 0, int herm = 0);
// Y (n x n complex) <- the matrix whose uplo triangle is stored in A: mode
// CPLX_SYM symmetric, CPLX_HERM Hermitian, CPLX_TRI triangular, CPLX_UNIT unit
// triangular.
enum : int { CPLX_SYM = 0, CPLX_HERM = 1, CPLX_TRI = 2, CPLX_UNIT = 3 };
// Inverse of `batch` complex triangular blocks of order n <= 64 (strides in
// complex elements); the other triangle of each n x n inverse is zeroed.
template <typename T>
void cplx_trtri_leaf(hipStream_t s, int uplo, int unit, int n, const T* A, long lda, long sA, T* Tinv, long ldinv,
                     long sInv, int batch);
// R (2n x 2n real, lower triangle) <- the real form [rho(a_ij)], rho(z) =
// [re -im; im re], of the Hermitian block whose uplo triangle A stores.
template <typename T>
void cplx_realform(hipStream_t s, int uplo, int n, const T* A, long lda, T* R, long ldr);
// From the real forms rho(L) (R) and rho(inv(L)) (Rinv, same leading
// dimension): L to A's uplo triangle (Upper: U = L^H) and inv(L) to Tinv's
// (Upper: inv(U)).
template <typename T>
void cplx_fromreal(hipStream_t s, int uplo, int n, const T* R, long ldr, T* A, long lda, const T* Rinv, T* Tinv,
                   long ldinv);
template <typename T>
void cplx_expand(hipStream_t s, int mode, int uplo, int n, const T* A, long lda, T* Y, long ldy);

// X <- X + D (m x n blocks).
void add2d(hipStream_t s, int m, int n, const double* D, long ldd, double* X, long ldx);
// Z <- float(Y 2^-e_j), e_j the exponent of colmax[j] (colmax null: no
// scaling); from_f32 undoes it: Y <- double(Z) 2^e_j, or Y += that (acc).
void to_f32(hipStream_t s, int m, int n, const double* Y, long ldy, const double* colmax, float* Z, long ldz);
void from_f32(hipStream_t s, int m, int n, const float* Z, long ldz, const double* colmax, double* Y, long ldy,
              bool acc);

// R <- B - A X (A n x n, X and B n x nrhs) with the k products summed in
// short MFMA runs and the run sums added without rounding error (TwoSum):
// the refinement residual, about sqrt(n / 64) times more accurate than one
// FP64 GEMM.  work: resid_scratch_bytes(n, nrhs) bytes of device memory.
size_t resid_scratch_bytes(int n, int nrhs);
void resid_accurate(hipStream_t s, int n, int nrhs, const double* A, long lda, const double* X, long ldx,
                    const double* B, long ldb, double* R, long ldr, void* work);

}  // namespace xk
