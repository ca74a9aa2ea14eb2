// stage3d.hip -- stage-kernel instantiations for the Kuhn-tetrahedron mesh
// (trait physics: Advection, Euler with each Riemann solver, with and
// without the entropy fix), P = 1..5.
#include "stage.hpp"

namespace mxm {
namespace {
template <int P>
StageLaunch pick(int phys, int variant) {
  if (phys == 0) return &launch_stage<Advection<double>, 3, P>;
  if (phys == 1) {
    switch (variant) {
      case 0: return &launch_stage<Euler<double, 0>, 3, P>;
      case 1: return &launch_stage<Euler<double, 1>, 3, P>;
      case 2: return &launch_stage<Euler<double, 2>, 3, P>;
      case 3: return &launch_stage<Euler<double, 3>, 3, P>;
      case 5: return &launch_stage<Euler<double, 1, 1>, 3, P>;
      case 6: return &launch_stage<Euler<double, 2, 1>, 3, P>;
      case 7: return &launch_stage<Euler<double, 3, 1>, 3, P>;
    }
  }
  return nullptr;
}
}  // namespace

StageLaunch find_stage_3d(int phys, int variant, int P) {
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
