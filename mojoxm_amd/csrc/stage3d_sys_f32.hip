// stage3d_sys_f32.hip -- Family3DSystem stage kernels, FP32-MFMA form, P = 1..5.
#include "stage_families.hpp"

namespace mxm {
StageLaunch stage_3d_sys_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  return pick_order<Family3DSystem, float>(phys, variant, P, cfg, lim, wide);
}
}  // namespace mxm
