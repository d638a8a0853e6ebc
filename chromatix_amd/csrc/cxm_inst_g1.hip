// Kernel instantiations for FFT size group 1 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(1, CXM_SIZES_G1)

// Column-pass launch-shape variants of the 1024-point pass (selected with
// cxm_set_option(CXM_OPT_COL_VARIANT, v)): (E, C) = 1: (16, 8), 2: (16, 16),
// 3: (32, 16), 4: (32, 8).
namespace cxm {
int launch_col_variant_1024(int v, const ColArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_col_t<1024, 16, 8>(a, s);
    case 2: return launch_col_t<1024, 16, 16>(a, s);
    case 3: return launch_col_t<1024, 32, 16>(a, s);
    case 4: return launch_col_t<1024, 32, 8>(a, s);
    default: return -1;
  }
}
}  // namespace cxm
