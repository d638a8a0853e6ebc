// Pipeline kernel instantiations for FFT size group 3 (see cxm_sizes.h).
#include "cxm_pipe_kernels.h"
CXM_DEFINE_PIPE_GROUP(3, CXM_SIZES_G3)
