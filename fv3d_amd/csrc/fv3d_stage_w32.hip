// fv3d_stage_w32.hip -- FP64 instantiations of the fused stage kernel
// (fv3d_stage_impl.h) for 32-column tiles (two window rows per wave).
#include "fv3d_stage_impl.h"

namespace fv3d_hip {

template hipError_t launch_stage_wx<double, STAGE_PLAIN, 32>(const StageLaunch<double>&,
                                                             hipStream_t);

}  // namespace fv3d_hip
