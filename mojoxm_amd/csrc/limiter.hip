// limiter.hip -- Barth-Jespersen slope limiter with Venkatakrishnan
// smoothing, split into a per-element theta and an apply that the readers
// of the state perform.
//
//   cell means     sum_i w_i q_i per (element, component), and the
//                  smallest and largest node value per (element,
//                  component), computed by the stage kernel's epilogue
//                  from the values it stores (compute_cell_averages_kernel,
//                  src/solver.mojo:384-407; cell_mean_kernel_2d,
//                  src/local_mesh_2d_gpu.mojo:246-264)
//   theta          one theta per element from the face neighbours' means
//                  (3D: all components, boundary self-neighbour counted,
//                  src/solver.mojo:410-530; 2D: component 0, boundary
//                  faces skipped, src/local_mesh_2d_gpu_limiter.mojo:74-153)
//   apply          q = mean + theta (q - mean) when theta < 1, done by the
//                  next stage kernel as it loads the state (q_in tile,
//                  partner traces, q_a), or by bj_apply_kernel, which
//                  writes the limited state and resets theta to 1.
//
// The reference takes the minimum of the Venkatakrishnan factor
//   g(x) = (D^2 + 2 D x + eps^2) / (D^2 + 2 x^2 + D x + eps^2)
// over the nodes, x = |q_i - mean|, D the bound on the side of q_i.  For
// x <= D/2, g >= 1 (D x >= 2 x^2), and beyond D/2 g decreases (the
// numerator of g' is D^3 + D eps^2 - 4 D x^2 - 4 D^2 x - 4 eps^2 x, which
// is negative there and falls with x), so the minimum over the nodes with
// q_i above the mean is g at the largest value, and below the mean g at
// the smallest: theta needs only the two extrema, and no pass over the
// state.  g is evaluated on those two nodes' own differences, as the
// per-node loop would.
#include <hip/hip_runtime.h>

#include "limiter.hpp"

namespace mxm {
namespace {

// one thread per element slot; padding slots and elements of inactive
// instances keep their theta
template <int NC, bool ALL>
__global__ __launch_bounds__(256) void bj_theta_kernel(double *__restrict__ theta, const double *__restrict__ mean,
                                                       const double *__restrict__ ext, const int *__restrict__ conn,
                                                       int K, int NF, double eps2, const int *__restrict__ tile_count,
                                                       const int *__restrict__ tile_inst,
                                                       const int *__restrict__ inst_active, int tepb) {
  constexpr int NCL = ALL ? NC : 1;  // limited components
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= K) return;
  const int T = e / tepb;
  if (e - T * tepb >= tile_count[T]) return;
  if (inst_active && !inst_active[tile_inst[T]]) return;
  double th = 1.0;
#pragma unroll
  for (int c = 0; c < NCL; ++c) {
    const double bm = mean[(size_t)e * NC + c];
    double mn = bm, mx = bm;
    for (int lf = 0; lf < NF; ++lf) {
      const int p = conn[(size_t)e * NF + lf];
      if (!ALL && p == e) continue;  // 2D: boundary faces do not constrain
      const double v = mean[(size_t)p * NC + c];
      mn = v < mn ? v : mn;
      mx = v > mx ? v : mx;
    }
    const double pD = mx - bm < 0.0 ? 0.0 : mx - bm;
    const double nD = bm - mn < 0.0 ? 0.0 : bm - mn;
#pragma unroll
    for (int s = 0; s < 2; ++s) {  // the smallest and the largest node value
      const double d = ext[((size_t)e * NCL + c) * 2 + s] - bm;
      const double da = d >= 0.0 ? d : -d;
      if (da <= 1.0e-30) continue;
      const double D = d > 0.0 ? pD : nD;
      const double D2 = D * D, d2 = da * da, Dd = D * da;
      const double al = (D2 + 2.0 * Dd + eps2) / (D2 + 2.0 * d2 + Dd + eps2);
      th = al < th ? al : th;
    }
  }
  theta[e] = th;
}

// the limited state written in place, then theta = 1: a block holds whole
// elements, one thread per node, so every node has read theta before the
// element's first thread resets it
template <class R, int NC>
__global__ __launch_bounds__(256) void bj_apply_kernel(R *__restrict__ q, double *__restrict__ theta,
                                                       const double *__restrict__ mean, int K, int NP) {
  const int epb = 256 / NP;
  const int el = threadIdx.x / NP, n = threadIdx.x - el * NP;
  const int e = blockIdx.x * epb + el;
  const bool mine = el < epb && e < K;
  const double th = mine ? theta[e] : 1.0;
  __syncthreads();
  if (!mine || !(th < 1.0)) return;
  if (n == 0) theta[e] = 1.0;
  R *qn = q + ((size_t)e * NP + n) * NC;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double bm = mean[(size_t)e * NC + c];
    qn[c] = R(bm + th * (double(qn[c]) - bm));
  }
}

__global__ void fill_kernel(double *p, long long n, double v) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int NC>
void theta_nc(double *theta, const double *mean, const double *ext, const int *conn, int K, int NF, bool all,
              double eps2, hipStream_t s, const int *tile_count, const int *tile_inst, const int *inst_active,
              int tepb) {
  const dim3 grid((K + 255) / 256);
  if (all)
    hipLaunchKernelGGL((bj_theta_kernel<NC, true>), grid, dim3(256), 0, s, theta, mean, ext, conn, K, NF, eps2,
                       tile_count, tile_inst, inst_active, tepb);
  else
    hipLaunchKernelGGL((bj_theta_kernel<NC, false>), grid, dim3(256), 0, s, theta, mean, ext, conn, K, NF, eps2,
                       tile_count, tile_inst, inst_active, tepb);
}

template <class R, int NC>
void apply_nc(R *q, double *theta, const double *mean, int K, int NP, hipStream_t s) {
  const int epb = 256 / NP;
  hipLaunchKernelGGL((bj_apply_kernel<R, NC>), dim3((K + epb - 1) / epb), dim3(256), 0, s, q, theta, mean, K, NP);
}

}  // namespace

hipError_t launch_bj_theta(double *theta, const double *mean, const double *ext, const int *conn, int K, int NC,
                           int NF, int dim, double eps, hipStream_t s, const int *tile_count, int epb_tiles,
                           const int *tile_inst, const int *inst_active) {
  if (K <= 0) return hipSuccess;
  if (!tile_count || epb_tiles <= 0) return hipErrorInvalidValue;
  const bool all = dim == 3;
  const double e2 = eps * eps;
#define MXM_THETA(N) \
  case N: theta_nc<N>(theta, mean, ext, conn, K, NF, all, e2, s, tile_count, tile_inst, inst_active, epb_tiles); break;
  switch (NC) {
    MXM_THETA(1) MXM_THETA(3) MXM_THETA(4) MXM_THETA(5) MXM_THETA(6) MXM_THETA(7) MXM_THETA(9) MXM_THETA(17)
    default: return hipErrorInvalidValue;
  }
#undef MXM_THETA
  return hipGetLastError();
}

template <class R>
hipError_t launch_bj_apply(R *q, double *theta, const double *mean, int K, int NP, int NC, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  if (NP > 256) return hipErrorInvalidValue;
#define MXM_APPLY(N) \
  case N: apply_nc<R, N>(q, theta, mean, K, NP, s); break;
  switch (NC) {
    MXM_APPLY(1) MXM_APPLY(3) MXM_APPLY(4) MXM_APPLY(5) MXM_APPLY(6) MXM_APPLY(7) MXM_APPLY(9) MXM_APPLY(17)
    default: return hipErrorInvalidValue;
  }
#undef MXM_APPLY
  return hipGetLastError();
}

hipError_t launch_fill(double *p, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

template hipError_t launch_bj_apply<double>(double *, double *, const double *, int, int, int, hipStream_t);
template hipError_t launch_bj_apply<float>(float *, double *, const double *, int, int, int, hipStream_t);

}  // namespace mxm
