// Kernel instantiations for FFT size group 2 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(2, CXM_SIZES_G2)

// Column-pass launch-shape variants of the 2048-point pass: (E, C) =
// 1: (16, 8), 2: (32, 8), 3: (16, 2), 4: (32, 4).
namespace cxm {
int launch_col_variant_2048(int v, const ColArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_col_t<2048, 16, 8>(a, s);
    case 2: return launch_col_t<2048, 32, 8>(a, s);
    case 3: return launch_col_t<2048, 16, 2>(a, s);
    case 4: return launch_col_t<2048, 32, 4>(a, s);
    default: return -1;
  }
}
int col_variant_width_2048(int v) {
  static const int c[5] = {0, 8, 8, 2, 4};
  return (v > 0 && v < 5) ? c[v] : -1;
}
}  // namespace cxm
