// Kernel instantiations for FFT size group 1 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(1, CXM_SIZES_G1)

// Launch-shape variants of the 1024-point passes (cxm_set_option):
//   CXM_OPT_COL_VARIANT (E, C) = 1: (16, 8), 2: (8, 4), 3: (8, 8), 4: (4, 4),
//                                5: (4, 2), 6: (8, 2), 7: (16, 2)
//   CXM_OPT_ROW_VARIANT E      = 1: 8, 2: 4
namespace cxm {
int launch_col_variant_1024(int v, const ColArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_col_t<1024, 16, 8>(a, s);
    case 2: return launch_col_t<1024, 8, 4>(a, s);
    case 3: return launch_col_t<1024, 8, 8>(a, s);
    case 4: return launch_col_t<1024, 4, 4>(a, s);
    case 5: return launch_col_t<1024, 4, 2>(a, s);
    case 6: return launch_col_t<1024, 8, 2>(a, s);
    case 7: return launch_col_t<1024, 16, 2>(a, s);
    default: return -1;
  }
}
int col_variant_width_1024(int v) {
  static const int c[8] = {0, 8, 4, 8, 4, 2, 2, 2};
  return (v > 0 && v < 8) ? c[v] : -1;
}
int launch_row_variant_1024(int v, int dir, const RowArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_row_t<1024, 8>(dir, a, s);
    case 2: return launch_row_t<1024, 4>(dir, a, s);
    default: return -1;
  }
}
}  // namespace cxm
