// stage.hpp -- the fused DG stage kernel for gfx950 and its launcher.
//
// One launch computes, for every element, the weak-form right-hand side
//   rhs_i = sum_k sum_j D_k[i][j] (invJ F(q_j))_k
//           - inv_vol * sum_f sum_r Lift_f[i][r] s_f area_f F*_{f,r}
//           + S(q_i)
// and the Shu-Osher combine q_out = a q_a + b q_b + c dt rhs followed by the
// physics' positivity fix -- rk_stage_kernel, src/solver.mojo:161-337, and
// (for 2D) the face-flux + vol_lift_combine_rk kernel pair of
// src/local_mesh_2d_gpu_<phys>.mojo fused into one pass.
//
// Work decomposition (256 threads = 4 wave64 per block, EPB elements per
// block):
//   1. each thread starts the loads of its first face item (job record,
//      face normal and area, partner node, and the partner's trace when the
//      partner lies outside the tile); meanwhile the block stages its
//      contiguous q_in tile, the tile's geometry and (when small) the
//      reference operators in LDS with coalesced loads;
//   2. per (element, node): the nodal flux, mapped to reference directions
//      by invJ, into LDS;
//      per face item: the numerical flux (boundary faces: against the
//      physics' ghost state), signed and scaled by the face area, into LDS.
//      A face whose two elements are both in the tile is evaluated once by
//      its side-0 element, which also stores the negated value in the
//      partner's slot; faces to elements outside the tile are evaluated by
//      each side, as the reference evaluates every face from both sides;
//      side 0 is always the left state;
//   3. per (element, node): the D and Lift contractions over LDS, source,
//      RK combine, limit_state, store.
// Blocks are mapped to tiles so that each of the 8 XCDs walks a contiguous
// range of tiles: face neighbours then mostly hit the XCD's own L2.
#pragma once

#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace mxm {

template <class R>
struct StageArgs {
  const R *q_in, *q_a, *q_b;
  R *q_out;
  R a, b, cdt;
  int K;
  const R *invJ;       // [K][ND*ND] rows = reference directions
  const R *inv_vol;    // [K] 1/(6V) or 1/(2A)
  const R *fnormal;    // [Nf][ND]
  const R *farea;      // [Nf]
  const int *perm;     // [nperm][NFP] partner node per reference face slot
  const R *Dt;         // [ND][NP(j)][NP(i)]  transposed D_ref
  const R *Lt;         // [NF][NFP(r)][NP(i)] transposed Lift_ref
  const int *fmap;     // [NF][NFP] element node of each reference face slot
  const int *job_off;  // [ntiles+1] face-flux jobs of each tile
  const int4 *jobs;    // {el | lf<<7 | pin<<9 | plf<<16 | pslot<<18, partner,
                       //  face id, side | bc<<1 | perm<<4}
  const int *pslot;    // [npslot][NFP] partner reference slot per slot
  int ntiles;          // tiles the job lists were built for
  int post_comp;       // component scaled after the combine (GLM psi), -1 none
  R post_scale;
};

// Elements per 256-thread block for NP nodes per element; the host builds
// the face job lists with the same tiling.
constexpr int stage_epb(int NP) { return 256 / NP; }

template <int ND, int P>
struct Shape {
  static constexpr int NP = ND == 3 ? (P + 1) * (P + 2) * (P + 3) / 6 : (P + 1) * (P + 2) / 2;
  static constexpr int NFP = ND == 3 ? (P + 1) * (P + 2) / 2 : P + 1;
  static constexpr int NF = ND + 1;
  static constexpr int BLOCK = 256;
  static constexpr int EPB = stage_epb(NP);
};

// Bijective map from the hardware block index (dispatched round-robin over
// the 8 XCDs) to a tile index such that XCD x owns a contiguous tile range.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, loc = bid >> 3;
  return x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
}

// One face item: slot r of a face job, with everything fetched from
// global memory before the flux is evaluated.
template <class R, int NC>
struct FaceItem {
  int packed, r, side, bc, pn, pl;
  R n[3], area;
  R qp[NC];  // partner trace when the partner is outside the tile
};

template <int ND, int NP, int NFP, int NC, class R>
__device__ __forceinline__ void load_face_item(const StageArgs<R> &A, int j0, int t, int e0, int nE,
                                               FaceItem<R, NC> &it) {
  const int jb = t / NFP;
  it.r = t - jb * NFP;
  const int4 jr = A.jobs[j0 + jb];
  it.packed = jr.x;
  it.side = jr.w & 1;
  it.bc = (jr.w >> 1) & 7;
#pragma unroll
  for (int d = 0; d < ND; ++d) it.n[d] = A.fnormal[(size_t)jr.z * ND + d];
  if constexpr (ND == 2) it.n[2] = R(0);
  it.area = A.farea[jr.z];
  it.pn = A.perm[(jr.w >> 4) * NFP + it.r];
  it.pl = jr.y - e0;
  const bool need = !(it.side == 0 && it.bc != 0);
  if (need && !(it.pl >= 0 && it.pl < nE)) {
    const R *g = A.q_in + ((size_t)jr.y * NP + it.pn) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) it.qp[c] = g[c];
  }
}

template <class Phys, int ND, int P, class R>
__global__ __launch_bounds__(256) void dg_stage_kernel(StageArgs<R> A, Phys phys) {
  using S = Shape<ND, P>;
  constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, EPB = S::EPB, NC = Phys::NC;
  constexpr int NOPS = ND * NP * NP + NF * NP * NFP;
  constexpr bool kOpsLDS = NOPS * (int)sizeof(R) <= 24 * 1024;
  __shared__ R sQ[EPB * NP * NC];
  __shared__ R sF[EPB * ND * NP * NC];  // [el][k][j][c]
  __shared__ R sS[EPB * NF * NFP * NC]; // [el][f][r][c]
  __shared__ R sJ[EPB * ND * ND + EPB]; // invJ then inv_vol
  __shared__ R sOp[kOpsLDS ? NOPS : 1];
  __shared__ int sMap[NF * NFP];

  const int tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int e0 = tile * EPB;
  const int nE = min(EPB, A.K - e0);
  if (nE <= 0) return;
  const int j0 = A.job_off[tile];
  const int nitems = (A.job_off[tile + 1] - j0) * NFP;

  // ---- 1. first face item of this thread, then the tile ------------------
  FaceItem<R, NC> it;
  if (tid < nitems) load_face_item<ND, NP, NFP, NC>(A, j0, tid, e0, nE, it);
  {
    const R *src = A.q_in + (size_t)e0 * NP * NC;
    const int n = nE * NP * NC;
    for (int t = tid; t < n; t += S::BLOCK) sQ[t] = src[t];
    const R *gJ = A.invJ + (size_t)e0 * ND * ND;
    for (int t = tid; t < nE * ND * ND; t += S::BLOCK) sJ[t] = gJ[t];
    for (int t = tid; t < nE; t += S::BLOCK) sJ[EPB * ND * ND + t] = A.inv_vol[e0 + t];
    for (int t = tid; t < NF * NFP; t += S::BLOCK) sMap[t] = A.fmap[t];
    if constexpr (kOpsLDS) {
      for (int t = tid; t < ND * NP * NP; t += S::BLOCK) sOp[t] = A.Dt[t];
      for (int t = tid; t < NF * NP * NFP; t += S::BLOCK) sOp[ND * NP * NP + t] = A.Lt[t];
    }
  }
  __syncthreads();
  const R *Dop = kOpsLDS ? sOp : A.Dt;
  const R *Lop = kOpsLDS ? sOp + ND * NP * NP : A.Lt;

  // ---- 2a. contravariant nodal fluxes ------------------------------------
  for (int t = tid; t < nE * NP; t += S::BLOCK) {
    const int el = t / NP, j = t - el * NP;
    R F[ND * NC];
    phys.template flux<ND>(&sQ[(el * NP + j) * NC], F);
    const R *iJ = &sJ[el * ND * ND];
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R v = R(0);
#pragma unroll
        for (int d = 0; d < ND; ++d) v += iJ[k * ND + d] * F[d * NC + c];
        sF[((el * ND + k) * NP + j) * NC + c] = v;
      }
  }

  // ---- 2b. face fluxes ----------------------------------------------------
  for (int t = tid; t < nitems; t += S::BLOCK) {
    if (t != tid) load_face_item<ND, NP, NFP, NC>(A, j0, t, e0, nE, it);
    const int el = it.packed & 127, lf = (it.packed >> 7) & 3, pin = (it.packed >> 9) & 127;
    R qm[NC], qo[NC], ql[NC], qr[NC];
    const int mine = (el * NP + sMap[lf * NFP + it.r]) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) qm[c] = sQ[mine + c];
    if (it.pl >= 0 && it.pl < nE) {
      const int other = (it.pl * NP + it.pn) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) qo[c] = sQ[other + c];
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) qo[c] = it.qp[c];
    }
    // left = side-0 trace; right = side-1 trace or the ghost of the left
#pragma unroll
    for (int c = 0; c < NC; ++c) ql[c] = it.side == 0 ? qm[c] : qo[c];
    if (it.bc)
      phys.ghost(ql, it.bc, it.n, qr);
    else
#pragma unroll
      for (int c = 0; c < NC; ++c) qr[c] = it.side == 0 ? qo[c] : qm[c];
    R Fs[NC];
    phys.numerical_flux(ql, qr, it.n, Fs);
    const R s = it.side == 0 ? it.area : -it.area;
#pragma unroll
    for (int c = 0; c < NC; ++c) sS[((el * NF + lf) * NFP + it.r) * NC + c] = s * Fs[c];
    if (pin) {
      const int plf = (it.packed >> 16) & 3, rs = A.pslot[(it.packed >> 18) * NFP + it.r];
#pragma unroll
      for (int c = 0; c < NC; ++c) sS[(((pin - 1) * NF + plf) * NFP + rs) * NC + c] = -(s * Fs[c]);
    }
  }
  __syncthreads();

  // ---- 3. contractions, source, RK combine, limit, store ----------------
  for (int t = tid; t < nE * NP; t += S::BLOCK) {
    const int el = t / NP, i = t - el * NP;
    R vol[NC], fac[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) vol[c] = fac[c] = R(0);
    for (int k = 0; k < ND; ++k) {
      const R *Fk = &sF[(el * ND + k) * NP * NC];
      const R *Dk = Dop + k * NP * NP;
#pragma unroll 4
      for (int j = 0; j < NP; ++j) {
        const R dv = Dk[j * NP + i];
#pragma unroll
        for (int c = 0; c < NC; ++c) vol[c] += dv * Fk[j * NC + c];
      }
    }
    const R *Se = &sS[el * NF * NFP * NC];
#pragma unroll 4
    for (int fr = 0; fr < NF * NFP; ++fr) {
      const R lv = Lop[fr * NP + i];
#pragma unroll
      for (int c = 0; c < NC; ++c) fac[c] += lv * Se[fr * NC + c];
    }
    const R iv = sJ[EPB * ND * ND + el];
    const R *qi = &sQ[(el * NP + i) * NC];
    R src[NC];
    if constexpr (Phys::kSource)
      phys.source(qi, src);
    else
#pragma unroll
      for (int c = 0; c < NC; ++c) src[c] = R(0);
    const size_t g = ((size_t)(e0 + el) * NP + i) * NC;
    R out[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const R rhs = vol[c] - iv * fac[c] + src[c];
      out[c] = A.a * A.q_a[g + c] + A.b * A.q_b[g + c] + A.cdt * rhs;
    }
    if constexpr (Phys::kLimit) phys.limit(out);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c == A.post_comp) out[c] *= A.post_scale;
#pragma unroll
    for (int c = 0; c < NC; ++c) A.q_out[g + c] = out[c];
  }
}

template <class Phys, int ND, int P>
hipError_t launch_stage(const StageArgs<double> &a, const double *prm, hipStream_t s) {
  using S = Shape<ND, P>;
  const Phys ph = Phys::from(prm);
  const int nb = (a.K + S::EPB - 1) / S::EPB;
  if (nb <= 0) return hipSuccess;
  if (nb != a.ntiles) return hipErrorInvalidValue;  // job lists built for another tiling
  hipLaunchKernelGGL((dg_stage_kernel<Phys, ND, P, double>), dim3(nb), dim3(S::BLOCK), 0, s, a, ph);
  return hipGetLastError();
}

using StageLaunch = hipError_t (*)(const StageArgs<double> &, const double *, hipStream_t);

// Lookups implemented in the per-family translation units; nullptr when
// the (physics, variant, P) combination is not compiled.
StageLaunch find_stage_3d(int phys, int variant, int P);
StageLaunch find_stage_2d_trait(int phys, int variant, int P);
StageLaunch find_stage_2d_module(int phys, int variant, int P);

}  // namespace mxm
