// init.h -- host problem generators (InitFunctor, Init.h:228-279).
#pragma once

#include "functors.h"
#include "sim_info.h"

namespace fv3d {

// Q (Ntz, Nty, Ntx, 5) with the domain set for cfg.problem; ghosts zero
// (the device builds them from the boundary maps).  Throws
// std::runtime_error("Error unknown problem ...") as Init.h:246-247 does.
void initialize(const RunConfig& cfg, HostArray& Q);

}  // namespace fv3d
