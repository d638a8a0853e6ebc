// Kernel instantiations for FFT size group 4 (see cxm_sizes.h).
#include "cxm_prop2d_kernels.h"
CXM_DEFINE_GROUP(4, CXM_SIZES_G4)
