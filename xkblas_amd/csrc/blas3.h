// blas3.h — BLAS-3 algorithms on device-resident column-major matrices.
// These are the MI355X restatement of the reference tile algorithms
// (blas/zgemm.c, zsyrk.c, ztrsm.c, zpotrf.c): instead of one task per tile,
// each step of the algorithm is one launch over all tiles it touches.
#pragma once
#include "common.h"
#include <functional>

namespace xk {

// Device scratch handed out by the runtime: slot-indexed, grow-only buffers.
void* scratch(int slot, size_t bytes);

// While an arena is set, scratch(slot, bytes) returns base + off[slot]
// (bytes must not exceed cap[slot]) instead of a runtime buffer: the
// allocation-free entry points carve their workspace out of a fixed array.
void scratch_arena_set(char* base, const size_t* off, const size_t* cap, int nslot);
void scratch_arena_clear();

struct Exec {
  hipStream_t s;      // main compute stream
  hipStream_t s_diag; // side stream for latency-bound diagonal-block work
  int mode;           // xkblas_mode_math_t
  int nsplit;         // CRT moduli for split modes (0 = automatic)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // POTRF's stream-ordering events (created per call when null)
};

// Scratch slots of dev_potrf, and the bytes each needs for order n, block nb.
enum { WS_DIAG0 = 0, WS_DIAG1 = 1, WS_PANEL = 2, WS_INV = 3 };
void potrf_scratch_caps(int n, int nb, size_t elt, size_t cap[4]);

template <typename T>
void dev_gemm(const Exec& x, int transA, int transB, int m, int n, int k, T alpha, const T* A, long lda,
              const T* B, long ldb, T beta, T* C, long ldc);

template <typename T>
void dev_syrk(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, T beta, T* C,
              long ldc);

// op(A) X = alpha B (side Left) or X op(A) = alpha B (Right); B is m x n.
// nb is the diagonal block order (multiple of 64 recommended).
template <typename T>
void dev_trsm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, T alpha, const T* A, long lda,
              T* B, long ldb, int nb);

// Optional work of a caller woven into dev_potrf (all members may be empty):
//  * beside_first() is called once to enqueue work on x.s right after the
//    first diagonal block's factorization has gone to the side stream (it
//    runs beside it); it must not touch the first block column of A, which
//    must be in place before the call;
//  * after_panel(k, ik) is called on the host once per block k, in order, to
//    enqueue work on x.s once block column k of the factor is final (L below
//    the diagonal block: Lower; U right of it: Upper); ik = the inverse of
//    diagonal block k (leading dimension nb), valid for work so enqueued;
//  * keep_inv receives the inverses of the factor's diagonal blocks of order
//    keep_nb, one after another (keep_nb x keep_nb each, a shorter last one
//    included), zeros outside their triangle: the layout of a TRSM sweep's
//    inverses.  keep_nb must divide the block order nb (or nb >= n).
template <typename T>
struct PotrfHooks {
  std::function<void()> beside_first;
  std::function<void(int, const T*)> after_panel;
  T* keep_inv = nullptr;
  int keep_nb = 0;
};

// In-place blocked Cholesky with block order nb.  info (device int) receives
// the LAPACK info as an atomic min over failing orders (INT_MAX = success);
// the caller initialises it.
template <typename T>
void dev_potrf(const Exec& x, int uplo, int n, T* A, long lda, int nb, int* info,
               const PotrfHooks<T>* hooks = nullptr);
// trailing updates of dev_potrf: 1 paired (default, also for -1), 0 one per panel
void potrf_set_pairing(int v);

// C <- alpha op(A) op(B) + beta C on the uplo triangle of the n x n C (uplo
// neither Lower nor Upper: all of C).
template <typename T>
void dev_gemmt(const Exec& x, int uplo, int transA, int transB, int n, int k, T alpha, const T* A, long lda,
               const T* B, long ldb, T beta, T* C, long ldc);

// C <- alpha (op(A) op(B)^T + op(B) op(A)^T) + beta C on the uplo triangle
// (op = identity for trans NoTrans: A, B n x k; else transposed: k x n).
template <typename T>
void dev_syr2k(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, const T* B, long ldb,
               T beta, T* C, long ldc);

// C <- alpha A B + beta C (Left) or alpha B A + beta C (Right), A symmetric
// with its uplo triangle stored; C, B m x n.
template <typename T>
void dev_symm(const Exec& x, int side, int uplo, int m, int n, T alpha, const T* A, long lda, const T* B, long ldb,
              T beta, T* C, long ldc);

// B <- alpha op(A) B (Left) or alpha B op(A) (Right), A triangular; B m x n.
template <typename T>
void dev_trmm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, T alpha, const T* A, long lda,
              T* B, long ldb);

// Complex C <- alpha op(A) op(B) + beta C (T = double: Z, float: C; complex
// numbers interleaved (re, im), alpha/beta as (re, im) pairs, leading
// dimensions in complex elements) as one real GEMM of order (2m, n, 2k).
template <typename T>
void dev_cgemm(const Exec& x, int transA, int transB, int m, int n, int k, const T* alpha, const T* A, long lda,
               const T* B, long ldb, const T* beta, T* C, long ldc);

// Complex C <- alpha op(A) op(A)^H + beta C on the uplo triangle (alpha,
// beta real; trans NoTrans: A n x k, else ConjTrans: A k x n), real diagonal.
template <typename T>
void dev_cherk(const Exec& x, int uplo, int trans, int n, int k, T alpha, const T* A, long lda, T beta, T* C,
               long ldc);

// Complex C <- alpha A B + beta C (Left) or alpha B A + beta C (Right), A
// Hermitian (herm) or symmetric with its uplo triangle stored.
template <typename T>
void dev_chemm(const Exec& x, int herm, int side, int uplo, int m, int n, const T* alpha, const T* A, long lda,
               const T* B, long ldb, const T* beta, T* C, long ldc);

// Complex GEMM on the uplo triangle of the n x n C (0: all of C).
template <typename T>
void dev_cgemmt(const Exec& x, int uplo, int transA, int transB, int n, int k, const T* alpha, const T* A, long lda,
                const T* B, long ldb, const T* beta, T* C, long ldc);

// Complex SYR2K (herm 0: alpha (op(A) op(B)^T + op(B) op(A)^T), trans N/T) or
// HER2K (herm 1: alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H, trans N/C,
// real diagonal), plus beta C, on the uplo triangle.  Also SYRK with B = A.
template <typename T>
void dev_csyr2k(const Exec& x, int herm, int uplo, int trans, int n, int k, const T* alpha, const T* A, long lda,
                const T* B, long ldb, const T* beta, T* C, long ldc);

// Complex TRSM: op(A) X = alpha B (Left) or X op(A) = alpha B (Right), B m x
// n overwritten by X; nb = diagonal block order.
template <typename T>
void dev_ctrsm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, const T* alpha, const T* A,
               long lda, T* B, long ldb, int nb);

// Complex Cholesky in place (uplo triangle; A = L L^H or U^H U), block order
// nb; info as for dev_potrf.
template <typename T>
void dev_cpotrf(const Exec& x, int uplo, int n, T* A, long lda, int nb, int* info);

// Complex TRMM: B <- alpha op(A) B (Left) or alpha B op(A) (Right).
template <typename T>
void dev_ctrmm(const Exec& x, int side, int uplo, int transA, int diag, int m, int n, const T* alpha, const T* A,
               long lda, T* B, long ldb);

// Solution of A X = B (A SPD, uplo triangle stored; B, X n x nrhs) from a
// factor computed under `xl` (a reduced-precision split mode) and FP64
// iterative refinement under `x`: LAPACK dsposv's scheme.  Returns the
// LAPACK info of the factor finally used; *iter = corrections applied, or
// -3 (reduced factor failed) / -(itermax + 1) (no convergence) when the
// factor of mode `xf` was used instead.  Synchronises the stream.
int dev_posv_ir(const Exec& x, const Exec& xl, const Exec& xf, int uplo, int n, int nrhs, const double* A, long lda,
                const double* B, long ldb, double* X, long ldx, int nb, int itermax, int* iter);

// Inverse of the triangular n x n block A into Tinv (other triangle zeroed).
template <typename T>
void dev_trtri(const Exec& x, hipStream_t s, int uplo, int diag, int n, const T* A, long lda, T* Tinv, long ldinv,
               int wslot);

}  // namespace xk
