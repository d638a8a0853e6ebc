// Pipeline kernel instantiations for FFT size group 2 (see cxm_sizes.h).
#include "cxm_pipe_kernels.h"
CXM_DEFINE_PIPE_GROUP(2, CXM_SIZES_G2)
