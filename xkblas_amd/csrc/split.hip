// split.hip — FP64-accurate split-precision products (placeholder until the
// CRT kernels land: the split modes fail loudly).
#include "split.h"

namespace xk {

template <typename T>
bool split_enabled(const Exec& x)
{
  return sizeof(T) == 8 && (x.mode == 2 || x.mode == 3);
}

template <typename T>
void split_gemm(const Exec&, int, int, int, int, int, T, const T*, long, const T*, long, T, T*, long, int)
{
  fatal("split_gemm", __FILE__, __LINE__, "split-precision mode is not built in this library version");
}

template bool split_enabled<double>(const Exec&);
template bool split_enabled<float>(const Exec&);
template void split_gemm<double>(const Exec&, int, int, int, int, int, double, const double*, long, const double*,
                                 long, double, double*, long, int);
template void split_gemm<float>(const Exec&, int, int, int, int, int, float, const float*, long, const float*, long,
                                float, float*, long, int);

}  // namespace xk
