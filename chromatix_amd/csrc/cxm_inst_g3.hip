// Kernel instantiations for FFT size group 3 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(3, CXM_SIZES_G3)
