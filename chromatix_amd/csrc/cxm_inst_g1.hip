// Kernel instantiations for FFT size group 1 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(1, CXM_SIZES_G1)
