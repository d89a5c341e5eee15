// fmm_runtime.hip -- libfmm_amd.so: the run-time-plan C-ABI (include/fmm_amd.h)
// over the gfx950 kernels of fmm_core.hpp.
#include "fmm_capi.hpp"
