// fv3d_adjoint_f32.hip -- FP32 instantiation of the adjoint stage kernels
// (fv3d_adjoint_impl.h).
#include "fv3d_adjoint_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_adjoint_stage<float>(const AdjLaunch<float>& L, hipStream_t s) {
  return launch_adjoint_stage_t<float>(L, s);
}

}  // namespace fv3d_hip
