// Kernel instantiations for FFT size group 3 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(3, CXM_SIZES_G3)

// Launch-shape variants of the 640-point column pass (cxm_set_option):
//   CXM_OPT_COL_VARIANT (E, C) = 1: (10, 4), 2: (10, 8), 3: (10, 2), 4: (20, 4)
namespace cxm {
int launch_col_variant_640(int v, const ColArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_col_t<640, 10, 4>(a, s);
    case 2: return launch_col_t<640, 10, 8>(a, s);
    case 3: return launch_col_t<640, 10, 2>(a, s);
    case 4: return launch_col_t<640, 20, 4>(a, s);
    default: return -1;
  }
}
int col_variant_width_640(int v) {
  static const int c[5] = {0, 4, 8, 2, 4};
  return (v > 0 && v < 5) ? c[v] : -1;
}
}  // namespace cxm
