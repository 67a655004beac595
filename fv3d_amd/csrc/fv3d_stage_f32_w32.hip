// fv3d_stage_f32_w32.hip -- FP32 instantiations of the fused stage kernel
// (fv3d_stage_impl.h) for 32-column tiles.  Built with the flags of
// fv3d_stage_f32.hip.
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

template hipError_t launch_stage_wx<float, STAGE_PLAIN, 32>(const StageLaunch<float>&, hipStream_t);
template hipError_t launch_stage_wx<float, STAGE_INC1, 32>(const StageLaunch<float>&, hipStream_t);
template hipError_t launch_stage_wx<float, STAGE_INC2, 32>(const StageLaunch<float>&, hipStream_t);

}  // namespace fv3d_hip
