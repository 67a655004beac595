// fv3d_batch.hip -- FP64 instantiations of the persistent multi-box stage
// kernel (fv3d_batch_impl.h), both tile widths, and the step-advance kernel.
#include "fv3d_batch_impl.h"

namespace fv3d_hip {

template <>
hipError_t launch_batch_stage<double>(const BatchLaunch& L, hipStream_t s) {
  if (L.wx == 32) return launch_batch_stage_wx<double, STAGE_PLAIN, 32>(L, s);
  return launch_batch_stage_wx<double, STAGE_PLAIN, 64>(L, s);
}

// end of a step: advance the step counter and clear the queue counters of
// both stages (a kernel, so that a captured step holds no memset node)
__global__ void batch_advance(long long* step, unsigned* counters, int n) {
  const int t = threadIdx.x;
  if (t < n) counters[t] = 0u;
  if (t == 0) *step += 1;
}

hipError_t launch_batch_advance(long long* step, unsigned* counters, int n, hipStream_t s) {
  hipLaunchKernelGGL(batch_advance, dim3(1), dim3(64), 0, s, step, counters, n);
  return hipGetLastError();
}

}  // namespace fv3d_hip
