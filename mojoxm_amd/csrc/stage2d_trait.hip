// stage2d_trait.hip -- stage-kernel instantiations of the trait physics
// (Advection, Euler with each Riemann solver; nz = 0) on the /-split
// triangle mesh, P = 1..5.
#include "stage.hpp"

namespace mxm {
namespace {
template <int P>
StageLaunch pick(int phys, int variant) {
  if (phys == 0) return &launch_stage<Advection<double>, 2, P>;
  if (phys == 1) {
    switch (variant) {
      case 0: return &launch_stage<Euler<double, 0>, 2, P>;
      case 1: return &launch_stage<Euler<double, 1>, 2, P>;
      case 2: return &launch_stage<Euler<double, 2>, 2, P>;
      case 3: return &launch_stage<Euler<double, 3>, 2, P>;
      case 5: return &launch_stage<Euler<double, 1, 1>, 2, P>;
      case 6: return &launch_stage<Euler<double, 2, 1>, 2, P>;
      case 7: return &launch_stage<Euler<double, 3, 1>, 2, P>;
    }
  }
  return nullptr;
}
}  // namespace

StageLaunch find_stage_2d_trait(int phys, int variant, int P) {
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
