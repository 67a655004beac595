// functors.h -- the reference's functor interface, over libfv3d_hip's C ABI.
//
// The reference builds UpdateFunctor, ComputeDtFunctor, BoundaryManager,
// ThermalConductionFunctor and ViscosityFunctor from Params and passes
// Kokkos Views between them (main.cpp:28-41).  Here the device state of a
// box lives in one fv3d_ctx (class Device).  Two forms of each functor keep
// their names and call semantics:
//  - on the Device's own state, the fused path (one kernel per RK stage),
//    which is what the driver runs;
//  - on DeviceArray, one reference Array (View<double****> of extents
//    (Ntz, Nty, Ntx, 5)) in device memory, one kernel per reference pass,
//    for callers that drive the passes one by one.
// Host arrays use the reference layout (Ntz, Nty, Ntx, 5); State is one
// cell's 5 values.  Errors throw std::runtime_error with the library's
// message (the C ABI itself never throws).
#pragma once

#include <array>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv3d_hip.h"

namespace fv3d {

using HostArray = std::vector<double>;
using State = std::array<double, FV3D_NFIELDS>;

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

// One Array of the reference in device memory, zero on creation like a new
// View.  It must not outlive its Device.
class DeviceArray {
 public:
  explicit DeviceArray(Device& d) : dev_(d) { check(fv3d_array_create(d.ctx(), &a_), d.ctx()); }
  DeviceArray(Device& d, const HostArray& host) : DeviceArray(d) { upload(host); }
  ~DeviceArray() { fv3d_array_destroy(a_); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  fv3d_array* get() const { return a_; }
  fv3d_ctx* ctx() const { return dev_.ctx(); }
  void upload(const HostArray& host) { check(fv3d_array_upload(a_, host.data()), ctx()); }
  void download(HostArray& host) const { check(fv3d_array_download(a_, host.data()), ctx()); }
  // Kokkos::deep_copy(*this, src)
  void copyFrom(const DeviceArray& src) { check(fv3d_array_copy(a_, src.a_), ctx()); }

 private:
  Device& dev_;
  fv3d_array* a_ = nullptr;
};

// ---- States.h: one cell ----------------------------------------------------
// primToCons / consToPrim of a State (States.h:20-45), evaluated on the device
inline State primToCons(const State& q, const fv3d_params& p) {
  State u;
  check(fv3d_prim_to_cons_states(p.gamma0, 1, q.data(), u.data()), nullptr);
  return u;
}
inline State consToPrim(const State& u, const fv3d_params& p) {
  State q;
  check(fv3d_cons_to_prim_states(p.gamma0, 1, u.data(), q.data()), nullptr);
  return q;
}

// ---- RiemannSolvers.h: hll / hllc on face-frame states (RiemannSolvers.h:6,39)
inline void hll(const State& qL, const State& qR, State& flux, double& pout, const fv3d_params& p) {
  check(fv3d_riemann(FV3D_HLL, p.gamma0, 1, qL.data(), qR.data(), flux.data(), &pout), nullptr);
}
inline void hllc(const State& qL, const State& qR, State& flux, double& pout, const fv3d_params& p) {
  check(fv3d_riemann(FV3D_HLLC, p.gamma0, 1, qL.data(), qR.data(), flux.data(), &pout), nullptr);
}

// ---- SimInfo.h: whole arrays -------------------------------------------------
// primToCons(Q, U) (SimInfo.h:323-331): Q (with ghosts) becomes the device state
inline void primToCons(Device& d, const HostArray& Q) { check(fv3d_set_prim(d.ctx(), Q.data(), 1), d.ctx()); }

// consToPrim(U, Q) (SimInfo.h:313-321), ghosts filled as inside update()
inline void consToPrim(Device& d, HostArray& Q) { check(fv3d_get_prim(d.ctx(), Q.data(), 1), d.ctx()); }

// the same two passes on device arrays, every cell, ghosts included
inline void primToCons(Device& d, const DeviceArray& Q, DeviceArray& U) {
  check(fv3d_prim_to_cons_array(d.ctx(), Q.get(), U.get()), d.ctx());
}
inline void consToPrim(Device& d, const DeviceArray& U, DeviceArray& Q) {
  check(fv3d_cons_to_prim_array(d.ctx(), U.get(), Q.get()), d.ctx());
}

// ComputeDt.h:9-64.  max_dt is accepted and unused, as in the reference.
class ComputeDtFunctor {
 public:
  explicit ComputeDtFunctor(const fv3d_params& p) : params_(p) {}
  // on the Device's state (the next step's dt, held on the device)
  double computeDt(Device& d, double /*max_dt*/, double t, bool diag) const {
    double inv[3], dt = 0.0;
    const int rc = fv3d_compute_dt(d.ctx(), inv, &dt);  // before dt is read below
    return finish(rc, d, inv, dt, t, diag);
  }
  // on a primitive array, as the reference calls it
  double computeDt(Device& d, const DeviceArray& Q, double /*max_dt*/, double t, bool diag) const {
    double inv[3], dt = 0.0;
    const int rc = fv3d_compute_dt_array(d.ctx(), Q.get(), inv, &dt);
    return finish(rc, d, inv, dt, t, diag);
  }

 private:
  double finish(int rc, Device& d, const double inv[3], double dt, double t, bool diag) const {
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
  fv3d_params params_;
  mutable bool warned_ = false;
};

// BoundaryConditions.h:137-215
class BoundaryManager {
 public:
  explicit BoundaryManager(const fv3d_params& p) : params_(p) {}
  // in place on a device primitive array
  void fillBoundaries(Device& d, DeviceArray& Q) const {
    check(fv3d_fill_boundaries_array(d.ctx(), Q.get()), d.ctx());
  }
  // on a host primitive array (stages it through the device)
  void fillBoundaries(HostArray& Q) const { check(fv3d_fill_boundaries(&params_, Q.data()), nullptr); }

 private:
  fv3d_params params_;
};

// ThermalConduction.h:24-98
class ThermalConductionFunctor {
 public:
  explicit ThermalConductionFunctor(const fv3d_params& p) : params_(p) {}
  void applyThermalConduction(Device& d, const DeviceArray& Q, DeviceArray& Unew, double dt) const {
    check(fv3d_apply_tc(d.ctx(), Q.get(), Unew.get(), dt), d.ctx());
  }

 private:
  fv3d_params params_;
};

// Viscosity.h:14-155
class ViscosityFunctor {
 public:
  explicit ViscosityFunctor(const fv3d_params& p) : params_(p) {}
  void applyViscosity(Device& d, const DeviceArray& Q, DeviceArray& Unew, double dt) const {
    check(fv3d_apply_visc(d.ctx(), Q.get(), Unew.get(), dt), d.ctx());
  }

 private:
  fv3d_params params_;
};

// Update.h:36-207.
class UpdateFunctor {
 public:
  explicit UpdateFunctor(const fv3d_params& p) : params_(p) {}

  // one Euler or SSP-RK2 step of the Device's state with the dt of the last
  // computeDt (held on the device).  Afterwards the device holds U^{n+1},
  // and stageState() gives what the reference's Q holds after update().
  void update(Device& d) const { check(fv3d_step(d.ctx()), d.ctx()); }
  // the same step with the caller's dt
  void update(Device& d, double dt) const { check(fv3d_step_dt(d.ctx(), dt), d.ctx()); }
  void stageState(Device& d, HostArray& Q) const {
    check(fv3d_get_stage_prim(d.ctx(), Q.data(), 1), d.ctx());
  }

  // update(Q, Unew, dt) on device arrays, pass by pass (mutates Q as the
  // reference does)
  void update(Device& d, DeviceArray& Q, DeviceArray& Unew, double dt) const {
    check(fv3d_update(d.ctx(), Q.get(), Unew.get(), dt), d.ctx());
  }
  void euler_step(Device& d, DeviceArray& Q, DeviceArray& Unew, double dt) const {
    check(fv3d_euler_step(d.ctx(), Q.get(), Unew.get(), dt), d.ctx());
  }
  // computeSlopes(Q) into this functor's slopesX/Y/Z (Update.h:49-51),
  // made on first use
  void computeSlopes(Device& d, const DeviceArray& Q) {
    if (!slopesX) computeSlopesStorage(d);
    check(fv3d_compute_slopes(d.ctx(), Q.get(), slopesX->get(), slopesY->get(), slopesZ->get()),
          d.ctx());
  }
  // computeFluxesAndUpdate(Q, Unew, dt) with the slopes of the last
  // computeSlopes (or zero slopes for PCM, which reads none)
  void computeFluxesAndUpdate(Device& d, const DeviceArray& Q, DeviceArray& Unew, double dt) {
    if (!slopesX) computeSlopesStorage(d);
    check(fv3d_fluxes_and_update(d.ctx(), Q.get(), slopesX->get(), slopesY->get(), slopesZ->get(),
                                 Unew.get(), dt),
          d.ctx());
  }

  std::unique_ptr<DeviceArray> slopesX, slopesY, slopesZ;

 private:
  void computeSlopesStorage(Device& d) {
    slopesX = std::make_unique<DeviceArray>(d);
    slopesY = std::make_unique<DeviceArray>(d);
    slopesZ = std::make_unique<DeviceArray>(d);
  }
  fv3d_params params_;
};

}  // namespace fv3d
