// Pipeline kernel instantiations for FFT size group 1 (see cxm_sizes.h).
#include "cxm_pipe_kernels.h"
CXM_DEFINE_PIPE_GROUP(1, CXM_SIZES_G1)

// Launch-shape variants of the 1024-point row pipeline (CXM_OPT_PIPE_ROW_VARIANT,
// or chosen automatically for launches of at most a few planes, where the
// default shape leaves most of the GPU idle): E = 1: 8, 2: 4.
namespace cxm {
int launch_pipe_row_variant_1024(int v, const PipeArgs& a, hipStream_t s) {
  switch (v) {
    case 1: return launch_pipe_row_t<1024, 8>(a, s);
    case 2: return launch_pipe_row_t<1024, 4>(a, s);
    default: return -1;
  }
}
}  // namespace cxm
