// stage2d_module.hip -- Family2DModule stage kernels, plain FP64 form, P = 1..5.
#include "stage_families.hpp"

namespace mxm {
StageLaunch stage_2d_module_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  return pick_order<Family2DModule, double>(phys, variant, P, cfg, lim, wide);
}
}  // namespace mxm
