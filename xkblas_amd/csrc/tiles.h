// tiles.h — workgroup -> tile order of the triangular GEMM-shaped kernels.
#pragma once
#include "common.h"

namespace xk {

// Lower triangle of T x T tiles (i >= j), total = T (T + 1) / 2 workgroups.
// Workgroup ids are dealt round-robin to the 8 XCDs: the id is remapped so
// that each XCD gets a contiguous range, and the range is walked in groups of
// GM tile rows, column by column inside a group, so the workgroups of one XCD
// that run together share the row panels of their group and a few column
// panels in its L2.  Group g (full) holds GM (g GM) + GM (GM + 1) / 2 tiles.
template <int GM>
__device__ __forceinline__ void tri_tile(int pid, int total, int T, int& i, int& j)
{
  if ((total & 7) == 0) pid = (pid & 7) * (total >> 3) + (pid >> 3);
  auto before = [](int h) { return (GM * GM * h * (h - 1)) / 2 + h * (GM * (GM + 1)) / 2; };
  // before(g) = a g^2 + b g: first guess from the quadratic, then exact
  const float a = 0.5f * GM * GM, b = 0.5f * GM * (GM + 1) - a;
  int g = (int)((-b + sqrtf(b * b + 4.0f * a * (float)pid)) / (2.0f * a));
  while (g > 0 && before(g) > pid) --g;
  while (before(g + 1) <= pid) ++g;
  const int r = pid - before(g);
  const int r0 = g * GM;           // first row of the group
  const int gr = min(GM, T - r0);  // rows of the group (the last may be short)
  if (r < r0 * gr) {               // columns left of the group's diagonal block
    j = r / gr;
    i = r0 + r % gr;
    return;
  }
  int rr = r - r0 * gr, jj = 0;    // diagonal block: column jj has gr - jj tiles
  while (rr >= gr - jj) {
    rr -= gr - jj;
    ++jj;
  }
  j = r0 + jj;
  i = r0 + jj + rr;
}

}  // namespace xk
