// Step kernels and their launches for double walkers (one translation unit per precision).
#include "pvd_sim_step.hpp"

namespace pvd {
PVD_SIM_OPS(double)
}  // namespace pvd
