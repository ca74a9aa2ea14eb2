// diag.hpp -- arguments of the diagnostics kernels (diag.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace mxm {

constexpr int kDiagMaxNC = 17;                  // components of any physics
constexpr int kDiagSums = 2 * kDiagMaxNC + 2;   // per-tile sums: q_c, q_c^2, KE, enstrophy

template <class R>
struct DiagArgs {
  const R *q;             // [slots][NP][NC], tile-ordered state
  const double *invJ;     // [slots][ND*ND], rows = reference directions
  const double *inv_vol;  // [slots] 1/(6V) or 1/(2A)
  const double *w;        // [NP] node weights, sum 1
  const double *Dn;       // [ND][NP][NP] nodal differentiation, d phi_j / d r_k at node i
  const int *tile_count;  // [tiles] elements per tile
  int epb;                // element slots per tile (epb * NP <= 256)
  int tile0, ntiles;      // the tiles [tile0, tile0 + ntiles) of the arrays
  int NC;                 // components
  int nvel;               // velocity components after the density (0: no KE / enstrophy)
  double *part;           // [ntiles][kDiagSums][2] compensated tile partials
  double *pmax;           // [ntiles][kDiagMaxNC]
  double *out;            // [3*NC + 2]: int q_c, int q_c^2, max |q_c|, KE, enstrophy
};

// Enqueues both kernels on s (args: DiagArgs<double> when fp64, else
// DiagArgs<float>).
hipError_t launch_diagnostics(int dim, int P, bool fp64, const void *args, hipStream_t s);

// velocity components stored after the density, per physics id (0: none)
inline int diag_nvel(int phys) {
  switch (phys) {
    case 1:            // trait Euler (rho, m[3], E), also on triangles
    case 9: return 3;  // 3D ideal MHD (rho, m[3], E, B[3], psi)
    case 2:            // Euler2D (rho, m[2], E)
    case 3:            // shallow water (h, h u[2])
    case 7:            // 3D shallow water (h, h u[2])
    case 4: return 2;  // GLM-MHD (rho, m[2], B[2], E, psi)
  }
  return 0;
}

}  // namespace mxm
