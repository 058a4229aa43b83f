// split.h — FP64-accurate products from BF16 / FP8 residue planes.
#pragma once
#include "blas3.h"

namespace xk {

// True when x.mode asks for a split mode and T is double.
template <typename T>
bool split_enabled(const Exec& x);

// C <- alpha op(A) op(B) + beta C (tri: written triangle) computed from CRT
// residues of row/column-scaled integer images of op(A), op(B).
template <typename T>
void split_gemm(const Exec& x, int transA, int transB, int m, int n, int k, T alpha, const T* A, long lda,
                const T* B, long ldb, T beta, T* C, long ldc, int tri);

// Number of CRT moduli for inner dimension k (requested >= 2 wins; else the
// smallest count keeping >= 46 bits of every row's largest entry), those
// guaranteed bits (a row of k equal-magnitude entries; rows with spread
// magnitudes keep more), and the number of residue GEMMs (planes) it takes;
// fp8 selects the FP8 moduli set (single planes for moduli <= 33, Karatsuba
// triples above).
int split_choose_count(int k, int requested, bool fp8);
int split_side_bits(int k, int nsplit, bool fp8);
int split_plane_count(int k, int requested, bool fp8);

// Residue planes reused across split products: registers op(X) (rows x k;
// row_contig: op(X)(i, kk) = X[i + kk ld], else X[kk + i ld]) as unchanging
// until the next registration or clear; products whose op(A) rows / op(B)
// columns are runs of its rows (first row a multiple of 256) take their
// exponents and residue planes from a cache.  The two latest registrations
// stay valid.
void split_cache_register(const double* X, long ld, int rows, int k, int row_contig);
void split_cache_clear();

// BF16 residue GEMM kernel (256 x 256 tiles, direct-to-LDS loads): 0 = 8
// waves of 128 x 64 (two per SIMD), 1 = 4 waves of 128 x 128.
void split_set_kernel(int v);
int split_get_kernel();

}  // namespace xk
