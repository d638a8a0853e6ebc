// Kernel instantiations for FFT size group 0 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(0, CXM_SIZES_G0)
