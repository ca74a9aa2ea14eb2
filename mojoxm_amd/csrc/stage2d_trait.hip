// stage2d_trait.hip -- Family2DTrait stage kernels, plain FP64 form, P = 1..5.
#include "stage_families.hpp"

namespace mxm {
StageLaunch stage_2d_trait_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  return pick_order<Family2DTrait, double>(phys, variant, P, cfg, lim, wide);
}
}  // namespace mxm
