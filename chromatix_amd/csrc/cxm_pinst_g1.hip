// Pipeline kernel instantiations for FFT size group 1 (see cxm_sizes.h).
#include "cxm_pipe_kernels.h"
CXM_DEFINE_PIPE_GROUP(1, CXM_SIZES_G1)
