// fv3d_stage.hip -- FP64 instantiations of the fused stage kernel
// (fv3d_stage_impl.h) for 64-column tiles, and the FP64 dispatch.  FP64
// division and square root are correctly rounded, so results are the
// reference's bit for bit.  The 32-column tiles are built in
// fv3d_stage_w32.hip.
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

extern template hipError_t launch_stage_wx<double, STAGE_PLAIN, 32>(const StageLaunch<double>&,
                                                                    hipStream_t);
template hipError_t launch_stage_wx<double, STAGE_PLAIN, 64>(const StageLaunch<double>&,
                                                             hipStream_t);

template <>
hipError_t launch_stage<double>(const StageLaunch<double>& L, hipStream_t s) {
  if (L.P.twx == 32) return launch_stage_wx<double, STAGE_PLAIN, 32>(L, s);
  return launch_stage_wx<double, STAGE_PLAIN, 64>(L, s);
}

}  // namespace fv3d_hip
