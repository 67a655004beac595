// io_manager.h -- IOManager::saveSolution (IOManager.h:79-192).
#pragma once

#include "functors.h"
#include "sim_info.h"

namespace fv3d {

class IOManager {
 public:
  explicit IOManager(const RunConfig& cfg);
  // Q: primitive host array (Ntz, Nty, Ntx, 5); dt is unused, as in the reference
  void saveSolution(const HostArray& Q, int iteration, double t, double dt);

 private:
  RunConfig cfg_;
};

}  // namespace fv3d
