// fv3d_kernels.h -- host-side launch interface of the libfv3d_hip kernels.
#pragma once

#include <hip/hip_runtime.h>
#include "fv3d_device.h"

namespace fv3d_hip {

// Stage tile shapes: WX columns (64: one window row per wave, or 32: two
// rows per wave) by stage_ty rows.  Without viscosity: 64 x 8 or 32 x 16
// (12 or 10 waves, a 168-VGPR budget).  With viscosity, FP32 keeps those
// shapes; FP64 takes 64 x 4 or 32 x 12 (8 waves, a 256-VGPR budget), where
// 64 x 8 spills about 80 VGPRs and runs 27% slower on the stratified box.
// Spilled VGPRs of the lean single-box kernels: FP32 none.  FP64 without
// conduction or viscosity: none for HLLC and 2 for HLL at 64 x 8 (none and
// 11-13 at 32 x 16); 8-31 with conduction alone; none with viscosity alone;
// 4 (HLLC) or 6 (HLL) with viscosity and conduction at 64 x 4, 4 at 32 x 12.
template <typename real, bool VISC, int WX>
constexpr int stage_ty_of() {
  return WX == 64 ? ((VISC && sizeof(real) == 8) ? 4 : 8) : ((VISC && sizeof(real) == 8) ? 12 : 16);
}
inline int stage_ty(bool visc, bool fp64, int wx) {
  return wx == 64 ? ((visc && fp64) ? 4 : 8) : ((visc && fp64) ? 12 : 16);
}

// Stage kernel variants.  STAGE_VAR_GENERAL tests the optional features at
// run time.  STAGE_VAR_C91 adds the precomputed C91 z ghosts.  The lean
// variants serve PLM runs without PCM_WB, well-balanced or C91 z-flux
// overrides or B02 conductivity (gravity stays a run-time switch) and drop
// those tests at compile time, so the kernel holds fewer scalar values.
// STAGE_VAR_LEAN also takes every cell size to be a power of two (x/h ==
// x*(1/h) exactly): about 30 fewer SGPRs spilled to VGPR lanes in the FP64
// blast kernel and a 4-5% shorter step, with the same bits.
// STAGE_VAR_LEAN_ANYH keeps the per-axis test (about 2% on the FP32
// ensemble, whose boxes mix power-of-two and other cell sizes).
enum { STAGE_VAR_GENERAL = 0, STAGE_VAR_C91 = 1, STAGE_VAR_LEAN = 2, STAGE_VAR_LEAN_ANYH = 3 };

// the kernel variant a box's parameters allow (see STAGE_VAR_*)
inline int stage_variant(const DevParams& P, int recon) {
  if (P.bcz == FV3D_BC_C91) return STAGE_VAR_C91;
  if (recon != FV3D_PLM || P.pcm_wb || P.wb_flux_zbc || P.c91_flux_zbc || P.kztab)
    return STAGE_VAR_GENERAL;
  return P.pow2x && P.pow2y && P.pow2z ? STAGE_VAR_LEAN : STAGE_VAR_LEAN_ANYH;
}

template <typename real>
struct StageLaunch {
  const real* Uin;
  const real* Urk;  // RK2 combine partner (U^n) or nullptr
  real* Uout;
  const unsigned long long* inv_cur;
  unsigned long long* inv_next;  // nullptr: no dt reduction
  unsigned long long* inv_zero;  // nullptr: nothing to clear
  int* health;
  int recon, riemann;
  bool tc, visc;
  int mode = 0;              // STAGE_PLAIN / STAGE_INC1 / STAGE_INC2 (FP32 only)
  real* Cio = nullptr;        // Kahan compensation, read and written (STAGE_INC2)
  DevParams P;
};

template <typename real>
hipError_t launch_stage(const StageLaunch<real>& L, hipStream_t s);
// the instantiations of one tile width (L.P.twx); launch_stage dispatches
template <typename real, int MODE, int WX>
hipError_t launch_stage_wx(const StageLaunch<real>& L, hipStream_t s);

// ---- many boxes: one persistent launch per stage (fv3d_batch_impl.h) ----
struct BoxDesc {
  DevParams P;
  long long off;             // element offset of the box in the pooled state buffers
  unsigned long long* ring;  // this box's inv_dt_hyp slots (3 used)
};

struct BatchBufs {
  void* A;     // state U^n (RK2 plain: updated in place by stage 2)
  void* B;     // U* (RK2, plain or compensated), or the other Euler state
  void* C;     // compensated RK2: the other U buffer
  void* comp;  // compensated RK2: Kahan compensation
  int* health;
};

struct BatchLaunch {
  const BoxDesc* boxes;
  const int2* items;      // (box, tile) work items, grouped by queue
  const int* qstart;      // 9 offsets: queue q owns items [qstart[q], qstart[q+1])
  unsigned* counters;     // 8 queue cursors of this stage
  const long long* step;  // device step counter
  BatchBufs bufs;
  int stage, ts, last_stage, grid;
  int recon, riemann;
  bool tc, visc;
  int mode;  // STAGE_PLAIN / STAGE_INC1 / STAGE_INC2
  int var = 0;  // stage kernel variant every box allows (STAGE_VAR_*)
  int wx = 64;  // tile width of every box (DevParams::twx)
};

template <typename real>
hipError_t launch_batch_stage(const BatchLaunch& L, hipStream_t s);
template <typename real, int MODE, int WX>
hipError_t launch_batch_stage_wx(const BatchLaunch& L, hipStream_t s);
hipError_t launch_batch_advance(long long* step, unsigned* counters, int n, hipStream_t s);

// ---- discrete adjoint of one stage (fv3d_adjoint_impl.h) ----
// out = sigma*v + dt J_R(U)^T (sigma*v) + alpha*extra, where the stage is
// S(U) = U + dt R(U) with dt frozen; grad[0..1] += (sigma*v) . dS/dkappa,
// (sigma*v) . dS/dmu.  Qb: [6][N] scratch, zero on entry and on return.
template <typename real>
struct AdjLaunch {
  const real* U;
  const real* v;
  real sigma;
  const real* extra;  // may be nullptr
  real alpha;
  real* Qb;
  real* out;
  double dt;
  double* grad;  // [2 + 2 * tiles]: dJ/dkappa, dJ/dmu, then per-tile partials (scratch)
  DevParams P;
  int recon, riemann;
  bool tc, visc;
  bool simple = false;   // one thread per face / cell with global atomics (check path)
  // tiled kernels: flush windows in colour passes with plain updates instead
  // of one launch with global atomics (measured slower on the 128^3 box:
  // each pass holds too few tiles to fill the GPU)
  bool colored = false;
  bool colored_diffusion = false;
  // finalize every cell in k_adj_finalize instead of most of them in the
  // z-face kernel (check path)
  bool no_fused_finalize = false;
};
template <typename real>
hipError_t launch_adjoint_stage(const AdjLaunch<real>& L, hipStream_t s);

// cons -> prim over the domain of U, written into a ghosted SoA array
// [5][Ntz][Nty][Ntx] (ghosts untouched)
template <typename real>
hipError_t launch_cons_to_prim_ghosted(const real* U, double* Q, const DevParams& P, int Ng,
                                       hipStream_t s);
// prim (SoA domain [5][Nz][Ny][Nx], double) -> cons state
template <typename real>
hipError_t launch_prim_to_cons(const double* Q, real* U, const DevParams& P, hipStream_t s);
// ghost fill of a ghosted SoA prim array: three ordered passes X, Y, Z
hipError_t launch_fill_boundaries(double* Q, const fv3d_params& p, hipStream_t s);
// max of inv_dt_hyp over the domain of U, into *inv (must be zeroed first)
template <typename real>
hipError_t launch_dt_reduce(const real* U, unsigned long long* inv, int* health,
                            const DevParams& P, hipStream_t s);
// batched Riemann solves over [n][5] states (face frame)
hipError_t launch_riemann_batch(int solver, double gamma, long n, const double* qL,
                                const double* qR, double* flux, double* pout, hipStream_t s);
// plain device copy of n doubles (n even)
hipError_t launch_copy(const double* a, double* b, long long n, hipStream_t s);
hipError_t launch_fill_zero_u64(unsigned long long* p, int n, hipStream_t s);
// *p = v, stream-ordered (a one-thread kernel, safe under graph capture)
hipError_t launch_put_u64(unsigned long long* p, unsigned long long v, hipStream_t s);

}  // namespace fv3d_hip
