// fv3d_batch_f32_w32.hip -- FP32 instantiations of the persistent multi-box
// stage kernel (fv3d_batch_impl.h) for 32-column tiles.  Built with the
// flags of fv3d_batch_f32.hip.
#include "fv3d_batch_impl.h"

namespace fv3d_hip {

template hipError_t launch_batch_stage_wx<float, STAGE_PLAIN, 32>(const BatchLaunch&, hipStream_t);
template hipError_t launch_batch_stage_wx<float, STAGE_INC1, 32>(const BatchLaunch&, hipStream_t);
template hipError_t launch_batch_stage_wx<float, STAGE_INC2, 32>(const BatchLaunch&, hipStream_t);

}  // namespace fv3d_hip
