// Kernel instantiations for FFT size group 2 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(2, CXM_SIZES_G2)
