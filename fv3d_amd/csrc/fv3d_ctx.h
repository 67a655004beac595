// fv3d_ctx.h -- the single-box context shared by the C-ABI translation
// units (fv3d_api.hip, fv3d_adjoint_api.hip).
#pragma once

#include <string>

#include "fv3d_hip.h"
#include "fv3d_host.h"
#include "fv3d_kernels.h"

struct fv3d_ctx {
  fv3d_params p;
  int precision = FV3D_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  void* A = nullptr;
  void* B = nullptr;
  void* C = nullptr;    // FP32 compensated RK2: U^{n+1} target
  void* Cmp = nullptr;  // FP32 compensated RK2: Kahan compensation
  size_t elem = sizeof(double);
  size_t state_bytes = 0;
  unsigned long long* ring = nullptr;  // 3 slots
  int* health = nullptr;
  double* qbuf = nullptr;  // ghosted SoA primitive staging, allocated on first use
  void* c91g = nullptr;    // boundary_z == C91: z-ghost planes of the stage input
  double* kztab = nullptr; // conductivity_mode == B02: face conductivities per z level
  long step = 0;
  bool have_state = false;
  bool have_stage = false;  // B holds the reference's post-update Q source
  fv3d_hip::DevParams dp;
  // the functor-level array path (fv3d_functors.hip), allocated on first
  // use: slopes x/y/z, RK2's U0 and Ustar (ghosted FP64 arrays), and its dt
  // accumulator
  void* fn_arrays[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  unsigned long long* fn_acc = nullptr;
  float adj_ms[3] = {0, 0, 0};  // last adjoint run: forward pass, backward pass, total
  int* adj_floor = nullptr;     // during an adjoint run: its frozen stages' density-floor flag
  size_t adj_store_bytes = 0;   // device memory for segments kept from the forward pass
  std::string err;
};

// records msg as the context's (and the library's) last error; returns code
int ctx_fail(fv3d_ctx* c, int code, const std::string& msg);
// recomputes the dt ring from the state in A and restarts the step count
int ctx_reset_dt(fv3d_ctx* c);

inline bool ctx_kahan(const fv3d_ctx* c) {
  return c->precision == FV3D_FP32 && c->p.time_stepping == FV3D_TS_RK2;
}

#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ctx_fail((ctx), FV3D_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
