// fv3d_batch_f32.hip -- FP32 instantiations of the persistent multi-box
// stage kernel (fv3d_batch_impl.h): plain stages and the two stages of the
// Kahan-compensated SSP-RK2.  Built with
// -fno-hip-fp32-correctly-rounded-divide-sqrt like fv3d_stage_f32.hip.
#include "fv3d_batch_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_batch_stage<float>(const BatchLaunch& L, hipStream_t s) {
  if (L.mode == STAGE_INC1) return batch_launch_mode<float, STAGE_INC1>(L, s);
  if (L.mode == STAGE_INC2) return batch_launch_mode<float, STAGE_INC2>(L, s);
  return batch_launch_mode<float, STAGE_PLAIN>(L, s);
}

}  // namespace fv3d_hip
