// stage2d_trait_f32.hip -- Family2DTrait stage kernels, FP32-MFMA form, P = 1..5.
#include "stage_families.hpp"

namespace mxm {
StageLaunch stage_2d_trait_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  return pick_order<Family2DTrait, float>(phys, variant, P, cfg, lim, wide);
}
}  // namespace mxm
