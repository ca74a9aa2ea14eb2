// stage3d_sys.hip -- Family3DSystem stage kernels, plain FP64 form, P = 1..5.
#include "stage_families.hpp"

namespace mxm {
StageLaunch stage_3d_sys_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  return pick_order<Family3DSystem, double>(phys, variant, P, cfg, lim, wide);
}
}  // namespace mxm
