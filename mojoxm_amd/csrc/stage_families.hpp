// stage_families.hpp -- the (physics, variant, P) -> stage-kernel tables.
// Each family is instantiated for one arithmetic form per translation unit
// (stage*.hip: FP64, stage*_f32.hip: FP32-MFMA) so they compile in
// parallel.  Euler variants: flux kind + 4 * entropy fix.
#pragma once

#include "stage.hpp"

namespace mxm {

// trait physics (Advection, Euler with each Riemann solver) on tetrahedra
struct Family3D {
  template <int P, class R, int W, bool L, bool WD>
  static StageLaunch pick(int phys, int variant) {
    if (phys == 0) return &launch_stage<Advection<R>, 3, P, R, W, L, WD>;
    if (phys == 1) {
      switch (variant) {
        case 0: return &launch_stage<Euler<R, 0>, 3, P, R, W, L, WD>;
        case 1: return &launch_stage<Euler<R, 1>, 3, P, R, W, L, WD>;
        case 2: return &launch_stage<Euler<R, 2>, 3, P, R, W, L, WD>;
        case 3: return &launch_stage<Euler<R, 3>, 3, P, R, W, L, WD>;
        case 5: return &launch_stage<Euler<R, 1, 1>, 3, P, R, W, L, WD>;
        case 6: return &launch_stage<Euler<R, 2, 1>, 3, P, R, W, L, WD>;
        case 7: return &launch_stage<Euler<R, 3, 1>, 3, P, R, W, L, WD>;
      }
    }
    return nullptr;
  }
};

// the same trait physics on triangles (nz = 0)
struct Family2DTrait {
  template <int P, class R, int W, bool L, bool WD>
  static StageLaunch pick(int phys, int variant) {
    if (phys == 0) return &launch_stage<Advection<R>, 2, P, R, W, L, WD>;
    if (phys == 1) {
      switch (variant) {
        case 0: return &launch_stage<Euler<R, 0>, 2, P, R, W, L, WD>;
        case 1: return &launch_stage<Euler<R, 1>, 2, P, R, W, L, WD>;
        case 2: return &launch_stage<Euler<R, 2>, 2, P, R, W, L, WD>;
        case 3: return &launch_stage<Euler<R, 3>, 2, P, R, W, L, WD>;
        case 5: return &launch_stage<Euler<R, 1, 1>, 2, P, R, W, L, WD>;
        case 6: return &launch_stage<Euler<R, 2, 1>, 2, P, R, W, L, WD>;
        case 7: return &launch_stage<Euler<R, 3, 1>, 2, P, R, W, L, WD>;
      }
    }
    return nullptr;
  }
};

// 2D module physics: Euler2D Rusanov/HLLC, SW Rusanov/HLL, GLM-MHD,
// Maxwell, Advection2D
struct Family2DModule {
  template <int P, class R, int W, bool L, bool WD>
  static StageLaunch pick(int phys, int variant) {
    switch (phys) {
      case 2: return variant ? &launch_stage<Euler2D<R, 1>, 2, P, R, W, L, WD> : &launch_stage<Euler2D<R, 0>, 2, P, R, W, L, WD>;
      case 3: return variant ? &launch_stage<SW2D<R, 1>, 2, P, R, W, L, WD> : &launch_stage<SW2D<R, 0>, 2, P, R, W, L, WD>;
      case 4: return &launch_stage<MHDGLM2D<R>, 2, P, R, W, L, WD>;
      case 5: return &launch_stage<Maxwell2D<R>, 2, P, R, W, L, WD>;
      case 6: return &launch_stage<Advection2D<R>, 2, P, R, W, L, WD>;
    }
    return nullptr;
  }
};

// the 3D systems with a Rusanov flux: shallow water, Maxwell, ideal MHD
// with GLM cleaning, five-moment two-fluid
struct Family3DSystem {
  template <int P, class R, int W, bool L, bool WD>
  static StageLaunch pick(int phys, int) {
    switch (phys) {
      case 7: return &launch_stage<ShallowWater3<R>, 3, P, R, W, L, WD>;
      case 8: return &launch_stage<Maxwell3<R>, 3, P, R, W, L, WD>;
      case 9: return &launch_stage<IdealMHD3<R>, 3, P, R, W, L, WD>;
      case 10: return &launch_stage<TwoFluid3<R>, 3, P, R, W, L, WD>;
    }
    return nullptr;
  }
};

template <class F, int P, class R, bool L, bool WD>
StageLaunch pick_cfg(int phys, int variant, int cfg) {
  switch (cfg) {
    case 0: return F::template pick<P, R, StageCfg<R>::waves[0], L, WD>(phys, variant);
    case 1: return F::template pick<P, R, StageCfg<R>::waves[1], L, WD>(phys, variant);
  }
  if constexpr (StageCfg<R>::count > 2)
    if (cfg == 2) return F::template pick<P, R, StageCfg<R>::waves[2], L, WD>(phys, variant);
  if constexpr (StageCfg<R>::count > 3)
    if (cfg == 3) return F::template pick<P, R, StageCfg<R>::waves[3], L, WD>(phys, variant);
  return nullptr;
}

// lim: the kernels with the limiter's epilogue and factor application
// (StageArgs::mean, ext, lim_*), instantiated beside the plain ones
template <class F, class R, bool L, bool WD>
StageLaunch pick_order_l(int phys, int variant, int P, int cfg) {
  switch (P) {
    case 1: return pick_cfg<F, 1, R, L, WD>(phys, variant, cfg);
    case 2: return pick_cfg<F, 2, R, L, WD>(phys, variant, cfg);
    case 3: return pick_cfg<F, 3, R, L, WD>(phys, variant, cfg);
    case 4: return pick_cfg<F, 4, R, L, WD>(phys, variant, cfg);
    case 5: return pick_cfg<F, 5, R, L, WD>(phys, variant, cfg);
  }
  return nullptr;
}
// wide: the kernels of the wide tile (stage_epb), FP32-MFMA form only; a
// shape without one returns its 256-thread kernels
template <class F, class R>
StageLaunch pick_order(int phys, int variant, int P, int cfg, bool lim, bool wide) {
  if constexpr (std::is_same<R, float>::value) {
    if (wide)
      return lim ? pick_order_l<F, R, true, true>(phys, variant, P, cfg)
                 : pick_order_l<F, R, false, true>(phys, variant, P, cfg);
  }
  return lim ? pick_order_l<F, R, true, false>(phys, variant, P, cfg)
             : pick_order_l<F, R, false, false>(phys, variant, P, cfg);
}

// one entry per (family, form), defined in the stage*.hip units
StageLaunch stage_3d_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_3d_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_3d_sys_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_3d_sys_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_2d_trait_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_2d_trait_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_2d_module_fp64(int phys, int variant, int P, int cfg, bool lim, bool wide);
StageLaunch stage_2d_module_fp32(int phys, int variant, int P, int cfg, bool lim, bool wide);

}  // namespace mxm
