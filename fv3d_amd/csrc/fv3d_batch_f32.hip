// fv3d_batch_f32.hip -- FP32 instantiations of the persistent multi-box
// stage kernel (fv3d_batch_impl.h) for 64-column tiles, and the FP32
// dispatch: plain stages and the two stages of the Kahan-compensated
// SSP-RK2.  Built with -fno-hip-fp32-correctly-rounded-divide-sqrt like
// fv3d_stage_f32.hip.  The 32-column tiles are built in
// fv3d_batch_f32_w32.hip.
#include "fv3d_batch_impl.h"

namespace fv3d_hip {

#define FV3D_BATCH_F32(MODE, WX) \
  template hipError_t launch_batch_stage_wx<float, MODE, WX>(const BatchLaunch&, hipStream_t)
extern FV3D_BATCH_F32(STAGE_PLAIN, 32);
extern FV3D_BATCH_F32(STAGE_INC1, 32);
extern FV3D_BATCH_F32(STAGE_INC2, 32);
FV3D_BATCH_F32(STAGE_PLAIN, 64);
FV3D_BATCH_F32(STAGE_INC1, 64);
FV3D_BATCH_F32(STAGE_INC2, 64);

template <int WX>
static hipError_t batch_f32(const BatchLaunch& L, hipStream_t s) {
  if (L.mode == STAGE_INC1) return launch_batch_stage_wx<float, STAGE_INC1, WX>(L, s);
  if (L.mode == STAGE_INC2) return launch_batch_stage_wx<float, STAGE_INC2, WX>(L, s);
  return launch_batch_stage_wx<float, STAGE_PLAIN, WX>(L, s);
}

template <>
hipError_t launch_batch_stage<float>(const BatchLaunch& L, hipStream_t s) {
  if (L.wx == 32) return batch_f32<32>(L, s);
  return batch_f32<64>(L, s);
}

}  // namespace fv3d_hip
