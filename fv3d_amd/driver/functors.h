// functors.h -- the reference's functor interface, over libfv3d_hip's C ABI.
//
// The reference builds UpdateFunctor, ComputeDtFunctor and BoundaryManager
// from Params and passes Kokkos Views between them (main.cpp:28-41).  Here
// the device state lives in one fv3d_ctx (class Device); the functors keep
// their names and call semantics and act on it.  Host arrays use the
// reference layout (Ntz, Nty, Ntx, 5).  Errors throw std::runtime_error
// with the library's message (the C ABI itself never throws).
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv3d_hip.h"

namespace fv3d {

using HostArray = std::vector<double>;

inline void check(int rc, const fv3d_ctx* ctx) {
  if (rc != FV3D_OK) throw std::runtime_error(std::string("fv3d: ") + fv3d_last_error(ctx));
}

class Device {
 public:
  Device(const fv3d_params& p, int precision, int device = 0) : params_(p) {
    check(fv3d_create(&p, precision, device, &ctx_), nullptr);
  }
  ~Device() { fv3d_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  fv3d_ctx* ctx() const { return ctx_; }
  const fv3d_params& params() const { return params_; }
  HostArray host_array() const {
    return HostArray((size_t)params_.Ntx * params_.Nty * params_.Ntz * FV3D_NFIELDS, 0.0);
  }

 private:
  fv3d_params params_;
  fv3d_ctx* ctx_ = nullptr;
};

// primToCons(Q, U) (SimInfo.h:323-331): Q (with ghosts) becomes the device state
inline void primToCons(Device& d, const HostArray& Q) { check(fv3d_set_prim(d.ctx(), Q.data(), 1), d.ctx()); }

// consToPrim(U, Q) (SimInfo.h:313-321), ghosts filled as inside update()
inline void consToPrim(Device& d, HostArray& Q) { check(fv3d_get_prim(d.ctx(), Q.data(), 1), d.ctx()); }

// ComputeDt.h:9-64.  max_dt is accepted and unused, as in the reference.
class ComputeDtFunctor {
 public:
  explicit ComputeDtFunctor(const fv3d_params& p) : params_(p) {}
  double computeDt(Device& d, double /*max_dt*/, double t, bool diag) const {
    double inv[3], dt;
    const int rc = fv3d_compute_dt(d.ctx(), inv, &dt);
    if (rc == FV3D_ENONFINITE) {
      // the reference leaves such cells out of its Max reduction and goes on
      if (!warned_)
        std::fprintf(stderr, "fv3d: warning at t=%g: %s; left out of dt, as in the reference\n", t,
                     fv3d_last_error(d.ctx()));
      warned_ = true;
    } else {
      check(rc, d.ctx());
    }
    if (diag)
      std::printf("Computing dts at (t=%g) : dt_hyp=%g; dt_TC=%g; dt_visc=%g\n", t, 1.0 / inv[0],
                  1.0 / inv[1], 1.0 / inv[2]);
    return dt;
  }

 private:
  fv3d_params params_;
  mutable bool warned_ = false;
};

// Update.h:36-207: one Euler or SSP-RK2 step with the dt of the last
// computeDt (held on the device).  Afterwards the device holds U^{n+1}, and
// stageState() gives what the reference's Q holds after update() returns.
class UpdateFunctor {
 public:
  explicit UpdateFunctor(const fv3d_params& p) : params_(p) {}
  void update(Device& d) const { check(fv3d_step(d.ctx()), d.ctx()); }
  void stageState(Device& d, HostArray& Q) const {
    check(fv3d_get_stage_prim(d.ctx(), Q.data(), 1), d.ctx());
  }

 private:
  fv3d_params params_;
};

// BoundaryConditions.h:137-215 on a host primitive array
class BoundaryManager {
 public:
  explicit BoundaryManager(const fv3d_params& p) : params_(p) {}
  void fillBoundaries(HostArray& Q) const { check(fv3d_fill_boundaries(&params_, Q.data()), nullptr); }

 private:
  fv3d_params params_;
};

}  // namespace fv3d
