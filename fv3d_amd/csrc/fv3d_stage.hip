// fv3d_stage.hip -- FP64 instantiations of the fused stage kernel
// (fv3d_stage_impl.h).  FP64 division and square root are correctly rounded,
// so results are the reference's bit for bit.
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_stage<double>(const StageLaunch<double>& L, hipStream_t s) {
  return launch_mode<double, STAGE_PLAIN>(L, s);
}

}  // namespace fv3d_hip
