// fv3d_stage_f32.hip -- FP32 instantiations of the fused stage kernel
// (fv3d_stage_impl.h) for 64-column tiles, and the FP32 dispatch: the plain
// stage and the two stages of the Kahan-compensated SSP-RK2.  This file is
// built with -fno-hip-fp32-correctly-rounded-divide-sqrt: FP32 division
// and square root use the hardware reciprocal/square-root (about 1 ulp),
// which the FP32 form's error budget against the FP64 path allows.  The
// 32-column tiles are built in fv3d_stage_f32_w32.hip.
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

#define FV3D_STAGE_F32(MODE, WX) \
  template hipError_t launch_stage_wx<float, MODE, WX>(const StageLaunch<float>&, hipStream_t)
extern FV3D_STAGE_F32(STAGE_PLAIN, 32);
extern FV3D_STAGE_F32(STAGE_INC1, 32);
extern FV3D_STAGE_F32(STAGE_INC2, 32);
FV3D_STAGE_F32(STAGE_PLAIN, 64);
FV3D_STAGE_F32(STAGE_INC1, 64);
FV3D_STAGE_F32(STAGE_INC2, 64);

template <int WX>
static hipError_t launch_f32(const StageLaunch<float>& L, hipStream_t s) {
  if (L.mode == STAGE_INC1) return launch_stage_wx<float, STAGE_INC1, WX>(L, s);
  if (L.mode == STAGE_INC2) return launch_stage_wx<float, STAGE_INC2, WX>(L, s);
  return launch_stage_wx<float, STAGE_PLAIN, WX>(L, s);
}

template <>
hipError_t launch_stage<float>(const StageLaunch<float>& L, hipStream_t s) {
  if (L.P.twx == 32) return launch_f32<32>(L, s);
  return launch_f32<64>(L, s);
}

}  // namespace fv3d_hip
