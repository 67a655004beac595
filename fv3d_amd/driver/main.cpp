// main.cpp -- the fv3d driver on libfv3d_hip: `fv3d file.ini [--precision
// fp64|fp32|fp32-plain] [--device N]`.
//
// Keeps the time loop of the reference's main (main.cpp:12-89): the save
// check before dt, dt from prim(U^n), the log line at step 0 and then every
// log_frequency+1 steps, t += dt with no clipping to the save times or tend,
// the loop test t + epsilon < tend, and the final snapshot taken from the
// state the reference's Q holds after the last update (prim(U*) for RK2).
// Extension: [run] max_steps stops the loop after that many steps.
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

#include "functors.h"
#include "init.h"
#include "io_manager.h"
#include "sim_info.h"

using namespace fv3d;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s file.ini [--precision fp64|fp32|fp32-plain] [--device N]\n", argv[0]);
    return 2;
  }
  int precision = FV3D_FP64, device = 0;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
      const std::string v = argv[++a];
      precision = v == "fp32" ? FV3D_FP32 : (v == "fp32-plain" ? FV3D_FP32_PLAIN : FV3D_FP64);
    } else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) {
      device = std::atoi(argv[++a]);
    }
  }
  try {
    RunConfig cfg = readInifile(argv[1]);
    for (const auto& w : cfg.warnings) std::fprintf(stderr, "fv3d warning: %s\n", w.c_str());
    const fv3d_params& params = cfg.params;

    HostArray Q;
    initialize(cfg, Q);  // InitFunctor::init (Init.h:253-278)
    Device dev(params, precision, device);
    UpdateFunctor update(params);
    ComputeDtFunctor computeDt(params);
    IOManager ioManager(cfg);
    primToCons(dev, Q);  // main.cpp:44

    double t = 0.0, dt = 0.0, next_save = 0.0;
    int ite = 0, next_log = 0;
    long steps = 0;
    while (t + params.epsilon < params.tend && (cfg.max_steps <= 0 || steps < cfg.max_steps)) {
      const bool save_needed = (t + params.epsilon > next_save);
      dt = computeDt.computeDt(dev, ite == 0 ? params.save_freq : next_save - t, t, next_log == 0);
      if (next_log == 0)
        next_log = params.log_frequency;
      else
        next_log--;
      if (save_needed) {
        std::printf(" - Saving at time %g\n", t);
        consToPrim(dev, Q);
        ioManager.saveSolution(Q, ite++, t, dt);
        next_save += params.save_freq;
      }
      update.update(dev);
      t += dt;
      ++steps;
    }
    // the stale Q of main.cpp:76; with no step taken it is still the initial Q
    if (steps > 0) update.stageState(dev, Q);
    ioManager.saveSolution(Q, ite++, t, dt);
    std::fflush(stdout);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
