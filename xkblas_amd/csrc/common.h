// common.h — shared definitions of libxkblas_amd (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace xk {

enum : int { NoTrans = 111, Trans = 112, ConjTrans = 113 };
enum : int { Upper = 121, Lower = 122 };
enum : int { NonUnit = 131, Unit = 132 };
enum : int { Left = 141, Right = 142 };

// Output-triangle selector of the GEMM kernel: FULL writes the whole block,
// LOWER (UPPER) writes only entries with row >= col (row <= col) of C.
enum : int { TRI_FULL = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

// Inner range of a GEMM tile when op(A) or op(B) is triangular (zeros
// outside): op(A) lower, rows i < m0 + BM need k < m0 + BM; op(A) upper,
// k >= m0; op(B) upper, columns j < n0 + BN need k < n0 + BN; op(B) lower,
// k >= n0.
enum : int { KRANGE_ALL = 0, KRANGE_A_LOWER = 1, KRANGE_A_UPPER = 2, KRANGE_B_UPPER = 3, KRANGE_B_LOWER = 4 };

[[noreturn]] void fatal(const char* what, const char* file, int line, const char* msg);

}  // namespace xk

#define XK_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) xk::fatal(#call, __FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

#define XK_CHECK(cond, msg)                                                  \
  do {                                                                       \
    if (!(cond)) xk::fatal(#cond, __FILE__, __LINE__, msg);                  \
  } while (0)
