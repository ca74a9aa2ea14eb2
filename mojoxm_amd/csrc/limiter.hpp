// limiter.hpp -- launcher of the Barth-Jespersen limiter kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace mxm {
constexpr int kLimMaxNC = 17;  // the widest state (two-fluid)
// q: state [K*NP*NC] (double or float) limited in place; the means, the
// bounds and theta are FP64 whatever the state precision.  mean: scratch
// [K*NC]; w: node weights [NP]; conn: element-centric connectivity
// (partner in .x).  In a batch, tile_inst[slot / epb] names the instance
// of a slot and elements of instances with inst_active == 0 are left
// alone (both null for a single solver).
template <class R>
hipError_t launch_limiter(R *q, double *mean, const double *w, const int4 *conn, int K, int NP, int NC,
                          int NF, int dim, double eps, hipStream_t s, const int *tile_inst = nullptr,
                          const int *inst_active = nullptr, int epb = 1);
}  // namespace mxm
