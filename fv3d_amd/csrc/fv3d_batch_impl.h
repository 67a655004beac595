// fv3d_batch_impl.h -- many independent boxes of uneven size advanced by one
// persistent kernel per RK stage.
//
// Work list: every (box, tile) pair, dealt to 8 queues -- one per XCD,
// boxes packed onto XCDs by cost (longest first) so that a box's tiles share
// one L2 -- and ordered longest first inside each queue.  A workgroup takes
// items from the queue of its own XCD (blocks are dealt to XCDs round-robin,
// block b -> XCD b%8) and steals from the others when it runs dry.
//
// The step counter lives on the device and the last kernel of a step
// (batch_advance) increments it and clears the queue counters, so one step
// -- stage kernels plus advance -- is a fixed launch sequence that a
// hipGraph replays unchanged.  Buffer parity and the per-box dt ring slots
// are derived from that counter inside the kernels.
#pragma once

#include "fv3d_stage_impl.h"

namespace fv3d_hip {

// which buffers a stage reads and writes, from the step parity
template <typename real, int MODE>
__device__ __forceinline__ void stage_pointers(const BatchBufs& bb, long long off, long long step,
                                               int stage, int ts, const real*& Uin,
                                               const real*& Urk, real*& Uout, real*& Cio) {
  real* A = (real*)bb.A + off;
  real* B = (real*)bb.B + off;
  Urk = nullptr;
  Cio = nullptr;
  if (MODE == STAGE_INC1 || MODE == STAGE_INC2) {
    real* C = (real*)bb.C + off;
    real* cur = (step & 1) ? C : A;
    real* nxt = (step & 1) ? A : C;
    if (MODE == STAGE_INC1) {  // U* into B
      Uin = cur;
      Uout = B;
    } else {  // from U* (B), compensated accumulation onto U^n (cur) into nxt
      Uin = B;
      Urk = cur;
      Cio = (real*)bb.comp + off;
      Uout = nxt;
    }
  } else if (ts == FV3D_TS_RK2) {
    if (stage == 0) { Uin = A; Uout = B; }
    else { Uin = B; Urk = A; Uout = A; }
  } else {
    Uin = (step & 1) ? B : A;
    Uout = (step & 1) ? A : B;
  }
}

template <typename real, int TY, int RECON, int RS, bool TC, bool VISC, int MODE, int VAR, int WX>
__global__ void __launch_bounds__(WX * (TY + 4))
batch_stage_kernel(const BoxDesc* __restrict__ boxes, const int2* __restrict__ items,
                   const int* __restrict__ qstart, unsigned* counters,
                   const long long* __restrict__ step_ptr, BatchBufs bb, int stage, int ts,
                   int last_stage) {
  __shared__ int s_item;
  const long long step = *step_ptr;
  const int q0 = blockIdx.x & 7;
  int tries = 0;  // used by thread 0 only
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      while (tries < 8) {
        const int q = (q0 + tries) & 7;
        const unsigned n = (unsigned)(qstart[q + 1] - qstart[q]);
        const unsigned got = atomicAdd(&counters[q], 1u);
        if (got < n) {
          it = qstart[q] + (int)got;
          break;
        }
        ++tries;
      }
      s_item = it;
    }
    __syncthreads();
    // uniform across the workgroup: lets the box descriptor load go through
    // scalar registers
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();  // every thread has read s_item before the next fetch
    if (item < 0) break;
    const int2 wi = items[item];
    const BoxDesc* bd = boxes + wi.x;
    const DevParams P = bd->P;
    const real *Uin, *Urk;
    real *Uout, *Cio;
    stage_pointers<real, MODE>(bb, bd->off, step, stage, ts, Uin, Urk, Uout, Cio);
    unsigned long long* ring = bd->ring;
    const unsigned long long* inv_cur = ring + (step % 3);
    unsigned long long* inv_next = last_stage ? ring + ((step + 1) % 3) : nullptr;
    unsigned long long* inv_zero = stage == 0 ? ring + ((step + 2) % 3) : nullptr;
    stage_tile<real, TY, RECON, RS, TC, VISC, MODE, VAR, WX>(Uin, Urk, Uout, Cio, inv_cur,
                                                             inv_next, inv_zero, bb.health, P,
                                                             wi.y);
  }
}

template <typename real, int TY, int RECON, int RS, bool TC, bool VISC, int MODE, int VAR, int WX>
static hipError_t batch_launch_t(const BatchLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((batch_stage_kernel<real, TY, RECON, RS, TC, VISC, MODE, VAR, WX>),
                     dim3(L.grid), dim3(WX * (TY + 4)), 0, s, L.boxes, L.items, L.qstart,
                     L.counters, L.step, L.bufs, L.stage, L.ts, L.last_stage);
  return hipGetLastError();
}

template <typename real, int RECON, int RS, int MODE, int VAR, int WX>
static hipError_t batch_launch_phys(const BatchLaunch& L, hipStream_t s) {
  constexpr int TV = stage_ty_of<real, true, WX>(), TH = stage_ty_of<real, false, WX>();
  if (L.tc && L.visc) return batch_launch_t<real, TV, RECON, RS, true, true, MODE, VAR, WX>(L, s);
  if (L.tc) return batch_launch_t<real, TH, RECON, RS, true, false, MODE, VAR, WX>(L, s);
  if (L.visc) return batch_launch_t<real, TV, RECON, RS, false, true, MODE, VAR, WX>(L, s);
  return batch_launch_t<real, TH, RECON, RS, false, false, MODE, VAR, WX>(L, s);
}

// L.var: the stage variant every box allows (the ensemble rejects C91);
// every box's tiles were sized for width WX (L.wx)
template <typename real, int MODE, int WX>
hipError_t launch_batch_stage_wx(const BatchLaunch& L, hipStream_t s) {
  if (L.wx != WX) return hipErrorInvalidValue;
  const bool plm = L.recon == FV3D_PLM, hll = L.riemann == FV3D_HLL;
  if (L.var == STAGE_VAR_LEAN) {
    if (hll) return batch_launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, STAGE_VAR_LEAN, WX>(L, s);
    return batch_launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, STAGE_VAR_LEAN, WX>(L, s);
  }
  if (L.var == STAGE_VAR_LEAN_ANYH) {
    if (hll) return batch_launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, STAGE_VAR_LEAN_ANYH, WX>(L, s);
    return batch_launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, STAGE_VAR_LEAN_ANYH, WX>(L, s);
  }
  constexpr int G = STAGE_VAR_GENERAL;
  if (plm && hll) return batch_launch_phys<real, FV3D_PLM, FV3D_HLL, MODE, G, WX>(L, s);
  if (plm) return batch_launch_phys<real, FV3D_PLM, FV3D_HLLC, MODE, G, WX>(L, s);
  if (hll) return batch_launch_phys<real, FV3D_PCM, FV3D_HLL, MODE, G, WX>(L, s);
  return batch_launch_phys<real, FV3D_PCM, FV3D_HLLC, MODE, G, WX>(L, s);
}

}  // namespace fv3d_hip
