// cfl.hip -- the largest signal speed over the nodes of a device-resident
// state, for a CFL time step dt = CFL h / (speed (2P + 1)).  The reference
// has no such reduction: its drivers estimate the speed a priori
// (examples/euler_vortex_2d_gpu.mojo:148-154) and the wave speeds its flux
// hooks return are discarded (src/solver.mojo:229,266,268).
//
// One 256-thread block per tile, each thread over the tile's element
// nodes tid, tid + 256, ...; each wave reduces with lane shuffles, the
// block over its 4 waves in LDS, and one block then reduces the tile
// maxima.  The speed is evaluated in
// FP64 from the state (FP64 or FP32).  Max is exact, so the result does
// not depend on the order; a NaN anywhere gives a NaN result.
#include <hip/hip_runtime.h>

#include "cfl.hpp"
#include "stage.hpp"

namespace mxm {
namespace {

__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ __forceinline__ double wave_nan_max(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
  return m;
}

template <class Phys, class S>
__global__ __launch_bounds__(256) void speed_tile_kernel(SpeedArgs<S> A) {
  __shared__ double sW[4];
  constexpr int NC = Phys::NC;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int T = A.tile0 + blockIdx.x;
  const int nE = A.tile_count[T];
  double sp = 0.0;
  const int inst = A.tile_inst ? A.tile_inst[T] : 0;
  const Phys phys = Phys::from(A.prm + (size_t)inst * kPrmMax);
  // a tile can hold more nodes than the block has threads (up to 512)
  for (int t = tid; t < nE * A.NP; t += 256) {
    const S *g = A.q + ((size_t)T * A.epb * A.NP + t) * NC;
    double q[NC];
    bool bad = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      q[c] = double(g[c]);
      bad = bad || q[c] != q[c];
    }
    // the physics' floors would replace a NaN density; report it instead
    sp = nan_max(sp, bad ? __builtin_nan("") : phys.max_speed(q));
  }
  sp = wave_nan_max(sp);
  if (lane == 0) sW[wave] = sp;
  __syncthreads();
  if (tid == 0) A.part[blockIdx.x] = nan_max(nan_max(sW[0], sW[1]), nan_max(sW[2], sW[3]));
}

__global__ __launch_bounds__(256) void speed_final_kernel(const double *part, int n, double *out) {
  __shared__ double sW[4];
  double m = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) m = nan_max(m, part[b]);
  m = wave_nan_max(m);
  if ((threadIdx.x & 63) == 0) sW[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = nan_max(nan_max(sW[0], sW[1]), nan_max(sW[2], sW[3]));
}

template <class Phys, class S>
hipError_t launch_ps(const SpeedArgs<S> &a, hipStream_t s) {
  if (a.ntiles > 0) hipLaunchKernelGGL((speed_tile_kernel<Phys, S>), dim3(a.ntiles), dim3(256), 0, s, a);
  hipLaunchKernelGGL(speed_final_kernel, dim3(1), dim3(256), 0, s, a.part, a.ntiles, a.out);
  return hipGetLastError();
}

// one functor per physics id (the flux variant does not change the speed)
template <class S>
hipError_t launch_s(int phys, const SpeedArgs<S> &a, hipStream_t s) {
  switch (phys) {
    case 0: return launch_ps<Advection<double>, S>(a, s);
    case 1: return launch_ps<Euler<double, 0>, S>(a, s);
    case 2: return launch_ps<Euler2D<double, 0>, S>(a, s);
    case 3: return launch_ps<SW2D<double, 0>, S>(a, s);
    case 4: return launch_ps<MHDGLM2D<double>, S>(a, s);
    case 5: return launch_ps<Maxwell2D<double>, S>(a, s);
    case 6: return launch_ps<Advection2D<double>, S>(a, s);
    case 7: return launch_ps<ShallowWater3<double>, S>(a, s);
    case 8: return launch_ps<Maxwell3<double>, S>(a, s);
    case 9: return launch_ps<IdealMHD3<double>, S>(a, s);
    case 10: return launch_ps<TwoFluid3<double>, S>(a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_max_speed(int phys, bool fp64, const void *args, hipStream_t s) {
  if (fp64) return launch_s<double>(phys, *static_cast<const SpeedArgs<double> *>(args), s);
  return launch_s<float>(phys, *static_cast<const SpeedArgs<float> *>(args), s);
}

}  // namespace mxm
