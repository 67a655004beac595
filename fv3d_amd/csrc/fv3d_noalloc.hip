// fv3d_noalloc.hip -- the allocation-free variant of fv3d_riemann, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight host loops, harnesses that forbid hipMalloc).
//
// fv3d_riemann stages its host face states and results in device buffers it
// allocates and frees on every call.  fv3d_riemann_noalloc stages them in
// `scratch`, a file-scope device array of FV3D_NOALLOC_MAX_FACES faces
// (16 doubles each: qL, qR, flux, pout; 8 MiB), which every device the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, so
// the device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and kept.  Every call on a device shares that
// device's array, so calls that overlap on one device race; nothing in it
// needs to start at zero (each call writes what it reads).
#include <mutex>
#include <string>

#include "fv3d_hip.h"
#include "fv3d_host.h"
#include "fv3d_kernels.h"

using namespace fv3d_hip;

namespace {

constexpr long kFaceDoubles = 16;  // qL[5], qR[5], flux[5], pout
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kFaceDoubles * FV3D_NOALLOC_MAX_FACES];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up on the
// first call there
int scratch_on_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return record_error(FV3D_EHIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return record_error(FV3D_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return record_error(FV3D_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return FV3D_OK;
}

}  // namespace

extern "C" {

int fv3d_riemann_noalloc(int solver, double gamma0, long n, const double* qL, const double* qR,
                         double* flux, double* pout) {
  if (n < 0 || (n > 0 && (!qL || !qR || !flux || !pout)))
    return record_error(FV3D_EINVAL, "bad argument");
  if (solver != FV3D_HLL && solver != FV3D_HLLC) return record_error(FV3D_EINVAL, "solver");
  if (n > FV3D_NOALLOC_MAX_FACES)
    return record_error(FV3D_EINVAL, "n = " + std::to_string(n) + " faces exceeds the " +
                                         std::to_string((long)FV3D_NOALLOC_MAX_FACES) +
                                         " of the allocation-free scratch");
  if (n == 0) return FV3D_OK;
  double* s = nullptr;
  if (int rc = scratch_on_current_device(&s)) return rc;
  double *dl = s, *dr = s + 5 * n, *df = s + 10 * n, *dp = s + 15 * n;
  const size_t b5 = (size_t)n * 5 * sizeof(double), b1 = (size_t)n * sizeof(double);
  // the null stream, synchronous copies: no stream or event is created either
  hipError_t e = hipMemcpy(dl, qL, b5, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dr, qR, b5, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_riemann_batch(solver, gamma0, n, dl, dr, df, dp, nullptr);
  if (e == hipSuccess) e = hipMemcpy(flux, df, b5, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pout, dp, b1, hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return record_error(FV3D_EHIP, std::string("riemann_noalloc: ") + hipGetErrorString(e));
  return FV3D_OK;
}

}  // extern "C"
