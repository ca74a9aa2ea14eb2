// cfl.hpp -- arguments of the largest-signal-speed reduction (cfl.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace mxm {

template <class S>
struct SpeedArgs {
  const S *q;             // [slots][NP][NC], tile-ordered state (FP64 or FP32)
  const double *prm;      // [instances][kPrmMax] physics parameters
  const int *tile_inst;   // [tiles] instance of each tile, or null (one instance)
  const int *tile_count;  // [tiles] elements per tile
  int epb, NP;            // element slots per tile (epb * NP <= 256), nodes per element
  int tile0, ntiles;      // the tiles [tile0, tile0 + ntiles) of the arrays
  double *part;           // [ntiles] tile maxima
  double *out;            // [1] the maximum (NaN if any node's speed is NaN)
};

// Enqueues both kernels on s (args: SpeedArgs<double> when fp64, else
// SpeedArgs<float>); the speed is evaluated in FP64 by the physics'
// max_speed.
hipError_t launch_max_speed(int phys, bool fp64, const void *args, hipStream_t s);

}  // namespace mxm
