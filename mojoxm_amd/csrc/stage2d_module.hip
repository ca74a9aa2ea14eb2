// stage2d_module.hip -- stage-kernel instantiations of the 2D module
// physics (Euler2D Rusanov/HLLC, SW Rusanov/HLL, GLM-MHD, Maxwell,
// Advection2D) on the /-split triangle mesh, P = 1..5.
#include "stage.hpp"

namespace mxm {
namespace {
template <int P>
StageLaunch pick(int phys, int variant) {
  switch (phys) {
    case 2: return variant ? &launch_stage<Euler2D<double, 1>, 2, P> : &launch_stage<Euler2D<double, 0>, 2, P>;
    case 3: return variant ? &launch_stage<SW2D<double, 1>, 2, P> : &launch_stage<SW2D<double, 0>, 2, P>;
    case 4: return &launch_stage<MHDGLM2D<double>, 2, P>;
    case 5: return &launch_stage<Maxwell2D<double>, 2, P>;
    case 6: return &launch_stage<Advection2D<double>, 2, P>;
  }
  return nullptr;
}
}  // namespace

StageLaunch find_stage_2d_module(int phys, int variant, int P) {
  switch (P) {
    case 1: return pick<1>(phys, variant);
    case 2: return pick<2>(phys, variant);
    case 3: return pick<3>(phys, variant);
    case 4: return pick<4>(phys, variant);
    case 5: return pick<5>(phys, variant);
  }
  return nullptr;
}
}  // namespace mxm
