// diag.hip -- domain diagnostics of a device-resident state: mass-weighted
// integrals of every component and of its square, max |q_c|, and for the
// physics with a velocity the kinetic energy and enstrophy, summed in FP64
// with compensated (TwoSum) arithmetic in a fixed order, so one state
// always gives the same bits.
//
// The reference's DiagnosticsWriter.record (src/diagnostics.mojo:161-223)
// downloads one component at a time and sums the raw nodal values on the
// host times domain_volume / total_dof (an unweighted nodal mean, :170).
// Here every element contributes V_e * sum_i w_i f(q_i) with the node
// weights w_i = int phi_i / |ref| (src/reference.mojo:650-659), exact for
// f of degree <= P, and the state never leaves the device.  KE =
// int 1/2 rho |u|^2 and enstrophy = int 1/2 rho |curl u|^2, with u = m/rho
// at the nodes and grad u from the nodal differentiation matrices mapped
// by invJ (the quantities the Taylor-Green driver approximates from
// momentum squares, examples/euler_taylor_green.mojo:181-195).
//
// Two kernels: one block per tile (one thread per element node: 256
// threads, 512 for tiles of more nodes) reduces its tile to a compensated
// pair (s, e) per sum and a max per component; one block then reduces the
// tile partials.
#include <hip/hip_runtime.h>

#include "diag.hpp"
#include "stage.hpp"

namespace mxm {
namespace {

// s + e == a + b exactly (Knuth's TwoSum; needs round-to-nearest and no
// reassociation, which this translation unit does not enable)
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}

// (s, e) += (s2, e2)
__device__ __forceinline__ void add_pair(double &s, double &e, double s2, double e2) {
  double t, err;
  two_sum(s, s2, t, err);
  s = t;
  e = (e + e2) + err;
}

// butterfly over the 64 lanes: every lane ends with the same pair
__device__ __forceinline__ void wave_sum(double &s, double &e) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double s2 = __shfl_xor(s, off, 64), e2 = __shfl_xor(e, off, 64);
    add_pair(s, e, s2, e2);
  }
}

__device__ __forceinline__ double wave_max(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  return m;
}

// B threads (256, or 512 for tiles of more than 256 nodes)
template <int ND, int P, class R, int B>
__global__ __launch_bounds__(B) void diag_tile_kernel(DiagArgs<R> A) {
  constexpr int NP = Shape<ND, P>::NP, NW = B / 64;
  constexpr double kRefVol = ND == 3 ? 1.0 / 6.0 : 0.5;
  __shared__ double sU[B * 3];
  __shared__ double sPart[NW][kDiagSums][2];
  __shared__ double sMax[NW][kDiagMaxNC];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int T = A.tile0 + blockIdx.x;
  const int nE = A.tile_count[T];
  const int NC = A.NC;
  const bool node = tid < nE * NP;
  const int el = tid / NP, i = tid - el * NP;
  const size_t slot = (size_t)T * A.epb + el;
  double qv[kDiagMaxNC];
  double vw = 0.0;  // V_e w_i
#pragma unroll
  for (int c = 0; c < kDiagMaxNC; ++c) qv[c] = 0.0;
  if (node) {
    const R *g = A.q + slot * NP * NC + (size_t)i * NC;
    for (int c = 0; c < NC; ++c) qv[c] = double(g[c]);
    vw = kRefVol / A.inv_vol[slot] * A.w[i];
  }
  if (A.nvel > 0 && node) {
    const double r = qv[0];
#pragma unroll
    for (int d = 0; d < 3; ++d) sU[tid * 3 + d] = d < A.nvel ? qv[1 + d] / r : 0.0;
  }
  __syncthreads();
  double ke = 0.0, ens = 0.0;
  if (A.nvel > 0 && node) {
    const double *iJ = A.invJ + slot * ND * ND;
    double g[3][3] = {{0.0}};  // g[a][b] = d u_a / d x_b
    for (int k = 0; k < ND; ++k) {
      const double *Dk = A.Dn + ((size_t)k * NP + i) * NP;
      double dr[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < NP; ++j) {
        const double d = Dk[j];
#pragma unroll
        for (int a = 0; a < 3; ++a) dr[a] += d * sU[(el * NP + j) * 3 + a];
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < ND; ++b) g[a][b] += iJ[k * ND + b] * dr[a];
    }
    const double w0 = g[2][1] - g[1][2], w1 = g[0][2] - g[2][0], w2 = g[1][0] - g[0][1];
    const double *u = &sU[tid * 3];
    const double r = qv[0];
    ke = vw * (0.5 * r * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    ens = vw * (0.5 * r * (w0 * w0 + w1 * w1 + w2 * w2));
  }
  // sums: [0, NC) int q_c, [NC, 2NC) int q_c^2, 2NC KE, 2NC + 1 enstrophy
  const int ns = 2 * NC + 2;
  for (int q = 0; q < ns; ++q) {
    double x;
    if (q < NC)
      x = vw * qv[q];
    else if (q < 2 * NC)
      x = vw * (qv[q - NC] * qv[q - NC]);
    else
      x = q == 2 * NC ? ke : ens;
    double s = x, e = 0.0;
    wave_sum(s, e);
    if (lane == 0) sPart[wave][q][0] = s, sPart[wave][q][1] = e;
  }
  for (int c = 0; c < NC; ++c) {
    const double m = wave_max(fabs(qv[c]));
    if (lane == 0) sMax[wave][c] = m;
  }
  __syncthreads();
  if (tid < ns) {
    double s = sPart[0][tid][0], e = sPart[0][tid][1];
    for (int w = 1; w < NW; ++w) add_pair(s, e, sPart[w][tid][0], sPart[w][tid][1]);
    A.part[((size_t)blockIdx.x * kDiagSums + tid) * 2] = s;
    A.part[((size_t)blockIdx.x * kDiagSums + tid) * 2 + 1] = e;
  } else if (tid >= 64 && tid < 64 + NC) {
    const int c = tid - 64;
    double m = sMax[0][c];
    for (int w = 1; w < NW; ++w) m = fmax(m, sMax[w][c]);
    A.pmax[(size_t)blockIdx.x * kDiagMaxNC + c] = m;
  }
}

// one block: the tile partials in a fixed order
template <class R>
__global__ __launch_bounds__(256) void diag_final_kernel(DiagArgs<R> A) {
  __shared__ double sPart[4][2];
  __shared__ double sMax[4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int NC = A.NC, ns = 2 * NC + 2, nb = A.ntiles;
  for (int q = 0; q < ns; ++q) {
    double s = 0.0, e = 0.0;
    for (int b = tid; b < nb; b += 256)
      add_pair(s, e, A.part[((size_t)b * kDiagSums + q) * 2], A.part[((size_t)b * kDiagSums + q) * 2 + 1]);
    wave_sum(s, e);
    if (lane == 0) sPart[wave][0] = s, sPart[wave][1] = e;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < 4; ++w) add_pair(s, e, sPart[w][0], sPart[w][1]);
      const int o = q < 2 * NC ? q : 3 * NC + (q - 2 * NC);
      A.out[o] = s + e;
    }
    __syncthreads();
  }
  for (int c = 0; c < NC; ++c) {
    double m = 0.0;
    for (int b = tid; b < nb; b += 256) m = fmax(m, A.pmax[(size_t)b * kDiagMaxNC + c]);
    m = wave_max(m);
    if (lane == 0) sMax[wave] = m;
    __syncthreads();
    if (tid == 0) A.out[2 * NC + c] = fmax(fmax(sMax[0], sMax[1]), fmax(sMax[2], sMax[3]));
    __syncthreads();
  }
}

template <int ND, int P, class R>
hipError_t launch_dp(const DiagArgs<R> &a, hipStream_t s) {
  const int nodes = a.epb * Shape<ND, P>::NP;
  if (nodes > 512) return hipErrorInvalidValue;
  if (nodes > 256)
    hipLaunchKernelGGL((diag_tile_kernel<ND, P, R, 512>), dim3(a.ntiles), dim3(512), 0, s, a);
  else
    hipLaunchKernelGGL((diag_tile_kernel<ND, P, R, 256>), dim3(a.ntiles), dim3(256), 0, s, a);
  hipLaunchKernelGGL((diag_final_kernel<R>), dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <class R>
hipError_t launch_r(int dim, int P, const DiagArgs<R> &a, hipStream_t s) {
  if (dim == 3) {
    switch (P) {
      case 1: return launch_dp<3, 1, R>(a, s);
      case 2: return launch_dp<3, 2, R>(a, s);
      case 3: return launch_dp<3, 3, R>(a, s);
      case 4: return launch_dp<3, 4, R>(a, s);
      case 5: return launch_dp<3, 5, R>(a, s);
    }
  } else {
    switch (P) {
      case 1: return launch_dp<2, 1, R>(a, s);
      case 2: return launch_dp<2, 2, R>(a, s);
      case 3: return launch_dp<2, 3, R>(a, s);
      case 4: return launch_dp<2, 4, R>(a, s);
      case 5: return launch_dp<2, 5, R>(a, s);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_diagnostics(int dim, int P, bool fp64, const void *args, hipStream_t s) {
  if (fp64) return launch_r<double>(dim, P, *static_cast<const DiagArgs<double> *>(args), s);
  return launch_r<float>(dim, P, *static_cast<const DiagArgs<float> *>(args), s);
}

}  // namespace mxm
