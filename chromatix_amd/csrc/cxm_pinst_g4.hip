// Pipeline kernel instantiations for FFT size group 4 (see cxm_sizes.h).
#include "cxm_pipe_kernels.h"
CXM_DEFINE_PIPE_GROUP(4, CXM_SIZES_G4)
