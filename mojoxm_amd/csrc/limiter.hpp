// limiter.hpp -- launcher of the Barth-Jespersen limiter kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace mxm {
constexpr int kLimMaxNC = 8;
// q: state [K*NP*NC] limited in place; mean: scratch [K*NC]; w: node
// weights [NP]; conn: element-centric connectivity (partner in .x).
hipError_t launch_limiter(double *q, double *mean, const double *w, const int4 *conn, int K, int NP,
                          int NC, int NF, int dim, double eps, hipStream_t s);
}  // namespace mxm
