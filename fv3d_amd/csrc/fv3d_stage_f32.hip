// fv3d_stage_f32.hip -- FP32 instantiations of the fused stage kernel
// (fv3d_stage_impl.h): the plain stage and the two stages of the
// Kahan-compensated SSP-RK2.  This file is built with
// -fno-hip-fp32-correctly-rounded-divide-sqrt: FP32 division and square
// root use the hardware reciprocal/square-root (about 1 ulp), which the
// FP32 form's error budget against the FP64 path allows.
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_stage<float>(const StageLaunch<float>& L, hipStream_t s) {
  if (L.mode == STAGE_INC1) return launch_mode<float, STAGE_INC1>(L, s);
  if (L.mode == STAGE_INC2) return launch_mode<float, STAGE_INC2>(L, s);
  return launch_mode<float, STAGE_PLAIN>(L, s);
}

}  // namespace fv3d_hip
