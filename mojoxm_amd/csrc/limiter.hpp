// limiter.hpp -- launchers of the Barth-Jespersen limiter's theta and
// apply kernels; the cell means and node extrema they read are written by
// the stage kernel's epilogue (StageArgs::mean, StageArgs::ext), and the
// next stage kernel applies theta as it loads the state
// (StageArgs::lim_*).
#pragma once
#include <hip/hip_runtime.h>

namespace mxm {
constexpr int kLimMaxNC = 17;  // the widest state (two-fluid)
// theta[K]: one factor per element slot from mean [K*NC] (sum_i w_i q_i),
// ext [K*NCL][2] (smallest, largest node value of the limited components:
// all NC in 3D, component 0 in 2D) and the neighbours' means through conn
// (the partner slot of each of an element's NF faces); FP64
// whatever the state precision.  tile_count[slot / epb]: real elements of
// each tile of epb slots (the others are padding and keep their theta).  In
// a batch, tile_inst[slot / epb] names the instance of a slot and elements
// of instances with inst_active == 0 keep theirs (both null for a single
// solver).
hipError_t launch_bj_theta(double *theta, const double *mean, const double *ext, const int *conn, int K, int NC,
                           int NF, int dim, double eps, hipStream_t s, const int *tile_count, int epb,
                           const int *tile_inst = nullptr, const int *inst_active = nullptr);
// q = mean + theta (q - mean) in place for every slot with theta < 1, then
// theta = 1 there: the state then holds the limited values and readers
// apply nothing more.
template <class R>
hipError_t launch_bj_apply(R *q, double *theta, const double *mean, int K, int NP, int NC, hipStream_t s);
// p[0..n) = v (a kernel, so that it can sit in a captured graph)
hipError_t launch_fill(double *p, long long n, double v, hipStream_t s);
}  // namespace mxm
