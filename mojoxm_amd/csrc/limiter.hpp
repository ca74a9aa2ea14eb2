// limiter.hpp -- launcher of the Barth-Jespersen limiter's theta + apply
// kernel; the cell means it reads are written by the stage kernel's
// epilogue (StageArgs::mean).
#pragma once
#include <hip/hip_runtime.h>

namespace mxm {
constexpr int kLimMaxNC = 17;  // the widest state (two-fluid)
// q: state [K*NP*NC] (double or float) limited in place; the means, the
// bounds and theta are FP64 whatever the state precision.  mean: cell
// means [K*NC] of q (sum_i w_i q_i); conn: element-centric connectivity
// (partner in .x); tile_count[slot / epb]: real elements of each tile of
// epb slots (the others are padding and left alone).  In a batch,
// tile_inst[slot / epb] names the instance of a slot and elements of
// instances with inst_active == 0 are left alone (both null for a single
// solver).
template <class R>
hipError_t launch_bj_apply(R *q, const double *mean, const int4 *conn, int K, int NP, int NC, int NF, int dim,
                           double eps, hipStream_t s, const int *tile_count, int epb,
                           const int *tile_inst = nullptr, const int *inst_active = nullptr);
}  // namespace mxm
