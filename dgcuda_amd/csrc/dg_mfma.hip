#include "dg_kernels.h"

namespace dgmi {

bool mfma_supported(int, int) { return false; }

hipError_t launch_stage_mfma(int, int, const StageArgs<float>&, dim3, hipStream_t) { return hipErrorInvalidValue; }

}  // namespace dgmi
