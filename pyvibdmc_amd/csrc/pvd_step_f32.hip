// Step kernels and their launches for float walkers (one translation unit per precision).
#include "pvd_sim_step.hpp"

namespace pvd {
PVD_SIM_OPS(float)
}  // namespace pvd
