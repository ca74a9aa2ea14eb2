// limiter.hip -- Barth-Jespersen slope limiter with Venkatakrishnan
// smoothing, run on q_out after a stage.
//
//   cell means     sum_i w_i q_i per (element, component), computed by the
//                  stage kernel's epilogue from the values it stores
//                  (compute_cell_averages_kernel, src/solver.mojo:384-407;
//                   cell_mean_kernel_2d, src/local_mesh_2d_gpu.mojo:246-264)
//   theta + apply  one theta per element from the face neighbours' means
//                  (3D: all components, boundary self-neighbour counted,
//                   src/solver.mojo:410-530; 2D: component 0, boundary
//                   faces skipped, src/local_mesh_2d_gpu_limiter.mojo:74-153)
//                  then q = mean + theta (q - mean) when theta < 1.
// The reference's separate theta and apply passes are one kernel here: a
// block holds whole elements, reduces theta over the element's nodes in
// LDS and applies it to the nodes it already holds in registers.  The
// kernel is instantiated per state width NC, so the per-component bounds
// live in registers.
#include <hip/hip_runtime.h>

#include "limiter.hpp"

namespace mxm {
namespace {

template <class R, int NC, bool ALL>
__global__ __launch_bounds__(256) void bj_theta_apply_kernel(R *__restrict__ q, const double *__restrict__ mean,
                                                             const int4 *__restrict__ conn, int K, int NP, int NF,
                                                             double eps2, const int *__restrict__ tile_count,
                                                             const int *__restrict__ tile_inst,
                                                             const int *__restrict__ inst_active, int tepb) {
  constexpr int NCL = ALL ? NC : 1;  // limited components
  __shared__ double sTheta[256];
  const int epb = 256 / NP;
  const int el = threadIdx.x / NP, n = threadIdx.x - el * NP;
  const int e = blockIdx.x * epb + el;
  // real elements only (a tile's first tile_count slots), of active instances
  const bool active = el < epb && e < K && e % tepb < tile_count[e / tepb] &&
                      !(inst_active && !inst_active[tile_inst[e / tepb]]);
  double theta = 1.0;
  double qv[NC], bm[NC];
  if (active) {
    const R *qn = q + ((size_t)e * NP + n) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      qv[c] = double(qn[c]);
      bm[c] = mean[(size_t)e * NC + c];
    }
    double mn[NCL], mx[NCL];
#pragma unroll
    for (int c = 0; c < NCL; ++c) mn[c] = mx[c] = bm[c];
    for (int lf = 0; lf < NF; ++lf) {
      const int p = conn[(size_t)e * NF + lf].x;
      if (!ALL && p == e) continue;  // 2D: boundary faces do not constrain
#pragma unroll
      for (int c = 0; c < NCL; ++c) {
        const double v = mean[(size_t)p * NC + c];
        mn[c] = v < mn[c] ? v : mn[c];
        mx[c] = v > mx[c] ? v : mx[c];
      }
    }
#pragma unroll
    for (int c = 0; c < NCL; ++c) {
      const double pD = mx[c] - bm[c] < 0.0 ? 0.0 : mx[c] - bm[c];
      const double nD = bm[c] - mn[c] < 0.0 ? 0.0 : bm[c] - mn[c];
      const double d = qv[c] - bm[c];
      const double da = d >= 0.0 ? d : -d;
      if (da <= 1.0e-30) continue;
      const double D = d > 0.0 ? pD : nD;
      const double D2 = D * D, d2 = da * da, Dd = D * da;
      const double al = (D2 + 2.0 * Dd + eps2) / (D2 + 2.0 * d2 + Dd + eps2);
      theta = al < theta ? al : theta;
    }
  }
  sTheta[threadIdx.x] = theta;
  __syncthreads();
  if (!active) return;
  double th = 1.0;
  for (int k = 0; k < NP; ++k) {
    const double v = sTheta[el * NP + k];
    th = v < th ? v : th;
  }
  if (!(th < 1.0)) return;
  R *qn = q + ((size_t)e * NP + n) * NC;
#pragma unroll
  for (int c = 0; c < NC; ++c) qn[c] = R(bm[c] + th * (qv[c] - bm[c]));
}

template <class R, int NC>
void launch_nc(R *q, const double *mean, const int4 *conn, int K, int NP, int NF, bool all, double eps2,
               hipStream_t s, const int *tile_count, const int *tile_inst, const int *inst_active, int tepb) {
  const int epb = 256 / NP;
  const dim3 grid((K + epb - 1) / epb);
  if (all)
    hipLaunchKernelGGL((bj_theta_apply_kernel<R, NC, true>), grid, dim3(256), 0, s, q, mean, conn, K, NP, NF, eps2,
                       tile_count, tile_inst, inst_active, tepb);
  else
    hipLaunchKernelGGL((bj_theta_apply_kernel<R, NC, false>), grid, dim3(256), 0, s, q, mean, conn, K, NP, NF, eps2,
                       tile_count, tile_inst, inst_active, tepb);
}

}  // namespace

template <class R>
hipError_t launch_bj_apply(R *q, const double *mean, const int4 *conn, int K, int NP, int NC, int NF, int dim,
                           double eps, hipStream_t s, const int *tile_count, int epb_tiles, const int *tile_inst,
                           const int *inst_active) {
  if (K <= 0) return hipSuccess;
  if (NP > 256 || !tile_count || epb_tiles <= 0) return hipErrorInvalidValue;
  const bool all = dim == 3;
  const double e2 = eps * eps;
  switch (NC) {
    case 1: launch_nc<R, 1>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 3: launch_nc<R, 3>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 4: launch_nc<R, 4>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 5: launch_nc<R, 5>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 6: launch_nc<R, 6>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 7: launch_nc<R, 7>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 9: launch_nc<R, 9>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    case 17: launch_nc<R, 17>(q, mean, conn, K, NP, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_bj_apply<double>(double *, const double *, const int4 *, int, int, int, int, int, double,
                                            hipStream_t, const int *, int, const int *, const int *);
template hipError_t launch_bj_apply<float>(float *, const double *, const int4 *, int, int, int, int, int, double,
                                           hipStream_t, const int *, int, const int *, const int *);

}  // namespace mxm
