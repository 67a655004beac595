// functor_check.cpp -- drives every functor facade of functors.h once, for
// the parity tests: `fv3d_functor_check in.bin out.bin`.
//
// in.bin:  fv3d_params (raw struct), dt (double), Q and U (host layout
//          (Ntz, Nty, Ntx, 5), doubles), qL and qR (one face-frame State each)
// out.bin: in this order, each a host array unless noted:
//   Q after BoundaryManager::fillBoundaries (device array, in place)
//   U2 = primToCons(Q) (arrays)
//   Q2 = consToPrim(U2) (arrays)
//   slopesX, slopesY, slopesZ of UpdateFunctor::computeSlopes(Q)
//   Unew = U after computeFluxesAndUpdate(Q, Unew, dt)
//   Unew after ThermalConductionFunctor::applyThermalConduction(Q, Unew, dt)
//   Unew after ViscosityFunctor::applyViscosity(Q, Unew, dt)
//   dt of ComputeDtFunctor::computeDt(Q)           (1 double)
//   primToCons(qL), consToPrim of that              (2 States)
//   hll(qL, qR) flux, pout; hllc(qL, qR) flux, pout (2 x (State, double))
//   Q3 and U3 after UpdateFunctor::update(Q3, U3, dt) from Q3 = Q, U3 = U
//   the state after UpdateFunctor::update(Device, dt) from primToCons(Q)
#include <cstdio>
#include <exception>

#include "functors.h"

using namespace fv3d;

namespace {

void read_exact(std::FILE* f, void* p, size_t n) {
  if (std::fread(p, 1, n, f) != n) throw std::runtime_error("short read");
}
void write_exact(std::FILE* f, const void* p, size_t n) {
  if (std::fwrite(p, 1, n, f) != n) throw std::runtime_error("short write");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  try {
    std::FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) throw std::runtime_error("cannot open input");
    fv3d_params p;
    double dt;
    read_exact(fi, &p, sizeof p);
    read_exact(fi, &dt, sizeof dt);
    Device dev(p, FV3D_FP64);
    HostArray Qh = dev.host_array(), Uh = dev.host_array(), out = dev.host_array();
    read_exact(fi, Qh.data(), Qh.size() * sizeof(double));
    read_exact(fi, Uh.data(), Uh.size() * sizeof(double));
    State qL, qR;
    read_exact(fi, qL.data(), sizeof qL);
    read_exact(fi, qR.data(), sizeof qR);
    std::fclose(fi);

    std::FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) throw std::runtime_error("cannot open output");
    auto put = [&](const DeviceArray& a) {
      a.download(out);
      write_exact(fo, out.data(), out.size() * sizeof(double));
    };

    DeviceArray Q(dev, Qh), U(dev, Uh), U2(dev), Q2(dev), Unew(dev, Uh);
    BoundaryManager bc(p);
    bc.fillBoundaries(dev, Q);
    put(Q);
    primToCons(dev, Q, U2);
    put(U2);
    consToPrim(dev, U2, Q2);
    put(Q2);
    UpdateFunctor update(p);
    update.computeSlopes(dev, Q);
    put(*update.slopesX);
    put(*update.slopesY);
    put(*update.slopesZ);
    update.computeFluxesAndUpdate(dev, Q, Unew, dt);
    put(Unew);
    ThermalConductionFunctor tc(p);
    tc.applyThermalConduction(dev, Q, Unew, dt);
    put(Unew);
    ViscosityFunctor visc(p);
    visc.applyViscosity(dev, Q, Unew, dt);
    put(Unew);
    ComputeDtFunctor cdt(p);
    const double dtq = cdt.computeDt(dev, Q, 0.0, 0.0, false);
    write_exact(fo, &dtq, sizeof dtq);

    const State uL = primToCons(qL, p);
    const State qback = consToPrim(uL, p);
    write_exact(fo, uL.data(), sizeof uL);
    write_exact(fo, qback.data(), sizeof qback);
    State f;
    double pout;
    hll(qL, qR, f, pout, p);
    write_exact(fo, f.data(), sizeof f);
    write_exact(fo, &pout, sizeof pout);
    hllc(qL, qR, f, pout, p);
    write_exact(fo, f.data(), sizeof f);
    write_exact(fo, &pout, sizeof pout);

    DeviceArray Q3(dev), U3(dev);
    Q3.copyFrom(Q);
    U3.copyFrom(U);
    update.update(dev, Q3, U3, dt);
    put(Q3);
    put(U3);

    primToCons(dev, Qh);  // the fused path: the Device's state from Q
    update.update(dev, dt);
    DeviceArray S(dev);
    check(fv3d_get_cons_array(dev.ctx(), S.get()), dev.ctx());
    put(S);
    std::fclose(fo);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
