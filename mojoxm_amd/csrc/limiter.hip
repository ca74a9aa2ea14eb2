// limiter.hip -- Barth-Jespersen slope limiter with Venkatakrishnan
// smoothing, run on q_out after a stage.
//
//   cell means     sum_i w_i q_i per (element, component)
//                  (compute_cell_averages_kernel, src/solver.mojo:384-407;
//                   cell_mean_kernel_2d, src/local_mesh_2d_gpu.mojo:246-264)
//   theta + apply  one theta per element from the face neighbours' means
//                  (3D: all components, boundary self-neighbour counted,
//                   src/solver.mojo:410-530; 2D: component 0, boundary
//                   faces skipped, src/local_mesh_2d_gpu_limiter.mojo:74-153)
//                  then q = mean + theta (q - mean) when theta < 1.
// The reference's separate theta and apply passes are one kernel here: a
// block holds whole elements, reduces theta over the element's nodes in
// LDS and applies it to the nodes it already holds in registers.
#include <hip/hip_runtime.h>

#include "limiter.hpp"

namespace mxm {

template <class R>
__global__ __launch_bounds__(256) void cell_mean_kernel(const R *__restrict__ q,
                                                        const double *__restrict__ w, int K, int NP,
                                                        int NC, double *__restrict__ mean) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * NC) return;
  const int e = t / NC, c = t - e * NC;
  const R *qe = q + (size_t)e * NP * NC + c;
  double s = 0.0;
  for (int n = 0; n < NP; ++n) s += w[n] * double(qe[n * NC]);
  mean[t] = s;
}

template <class R>
__global__ __launch_bounds__(256) void bj_theta_apply_kernel(R *__restrict__ q,
                                                             const double *__restrict__ mean,
                                                             const int4 *__restrict__ conn, int K,
                                                             int NP, int NC, int NF, int all_comps,
                                                             double eps2, const int *__restrict__ tile_inst,
                                                             const int *__restrict__ inst_active, int tepb) {
  __shared__ double sTheta[256];
  const int epb = 256 / NP;
  const int el = threadIdx.x / NP, n = threadIdx.x - el * NP;
  const int e = blockIdx.x * epb + el;
  const bool active = el < epb && e < K && !(inst_active && !inst_active[tile_inst[e / tepb]]);
  double theta = 1.0;
  double own[kLimMaxNC], mn[kLimMaxNC], mx[kLimMaxNC], qv[kLimMaxNC];
  const int ncl = all_comps ? NC : 1;
  if (active) {
    for (int c = 0; c < ncl; ++c) own[c] = mn[c] = mx[c] = mean[(size_t)e * NC + c];
    for (int lf = 0; lf < NF; ++lf) {
      const int p = conn[(size_t)e * NF + lf].x;
      if (!all_comps && p == e) continue;  // 2D: boundary faces do not constrain
      for (int c = 0; c < ncl; ++c) {
        const double v = mean[(size_t)p * NC + c];
        mn[c] = v < mn[c] ? v : mn[c];
        mx[c] = v > mx[c] ? v : mx[c];
      }
    }
    const R *qn = q + ((size_t)e * NP + n) * NC;
    for (int c = 0; c < NC; ++c) qv[c] = double(qn[c]);
    for (int c = 0; c < ncl; ++c) {
      const double pD = mx[c] - own[c] < 0.0 ? 0.0 : mx[c] - own[c];
      const double nD = own[c] - mn[c] < 0.0 ? 0.0 : own[c] - mn[c];
      const double d = qv[c] - own[c];
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
  for (int c = 0; c < NC; ++c) {
    const double bm = mean[(size_t)e * NC + c];
    qn[c] = R(bm + th * (qv[c] - bm));
  }
}

template <class R>
hipError_t launch_limiter(R *q, double *mean, const double *w, const int4 *conn, int K, int NP, int NC,
                          int NF, int dim, double eps, hipStream_t s, const int *tile_inst,
                          const int *inst_active, int epb_tiles) {
  if (K <= 0) return hipSuccess;
  if (NC > kLimMaxNC || NP > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(cell_mean_kernel<R>, dim3((K * NC + 255) / 256), dim3(256), 0, s, q, w, K, NP, NC, mean);
  const int epb = 256 / NP;
  hipLaunchKernelGGL(bj_theta_apply_kernel<R>, dim3((K + epb - 1) / epb), dim3(256), 0, s, q, mean, conn, K, NP,
                     NC, NF, dim == 3 ? 1 : 0, eps * eps, tile_inst, inst_active, epb_tiles);
  return hipGetLastError();
}

template hipError_t launch_limiter<double>(double *, double *, const double *, const int4 *, int, int, int, int,
                                           int, double, hipStream_t, const int *, const int *, int);
template hipError_t launch_limiter<float>(float *, double *, const double *, const int4 *, int, int, int, int,
                                          int, double, hipStream_t, const int *, const int *, int);

}  // namespace mxm
