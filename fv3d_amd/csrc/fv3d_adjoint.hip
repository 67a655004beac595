// fv3d_adjoint.hip -- FP64 instantiation of the adjoint stage kernels
// (fv3d_adjoint_impl.h).
#include "fv3d_adjoint_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_adjoint_stage<double>(const AdjLaunch<double>& L, hipStream_t s) {
  return launch_adjoint_stage_t<double>(L, s);
}

}  // namespace fv3d_hip
