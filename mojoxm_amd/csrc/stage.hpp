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

// Shu-Osher stage q_out = a q_a + b q_in + c dt L(q_in): in every SSPRK3
// stage of src/ssprk3.mojo:41-78 the b-term buffer is the stage input.
template <class R>
struct StageArgs {
  const R *q_in, *q_a;
  R *q_out;
  R a, b, cdt;
  int K;
  const R *invJ;       // [K][ND*ND] rows = reference directions
  const R *inv_vol;    // [K] 1/(6V) or 1/(2A)
  const R *fnormal;    // [Nf][ND]
  const R *farea;      // [Nf]
  const int *perm;     // [nperm][NFP] partner node per reference face slot
  const R *Dt;         // [ND*NP (k,j)][NPAD (i)] transposed D_ref, zero padded
  const R *Lt;         // [NF*NFP (f,r)][NPAD (i)] transposed Lift_ref, zero padded
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

// Contraction layout: the tile's (element, component) columns span
// ceil(EPB*NC/64) waves; the block's 4 waves form `groups` groups, each
// producing `IH` consecutive output nodes, so the operator entries a wave
// reads are wave-uniform.  The transposed operators are stored with
// groups*IH (>= NP) zero-padded node columns.
constexpr int stage_groups(int NP, int NC) {
  return (stage_epb(NP) * NC + 63) / 64 >= 4 ? 1 : 4 / ((stage_epb(NP) * NC + 63) / 64);
}
constexpr int stage_node_pad(int NP, int NC) {
  return stage_groups(NP, NC) * ((NP + stage_groups(NP, NC) - 1) / stage_groups(NP, NC));
}

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
  // Contraction layout: the (element, component) columns of the tile are
  // spread over one or more waves; the block's waves split into GROUPS
  // groups that each produce a contiguous range of IH output nodes, so the
  // operator entries a wave needs are wave-uniform (scalar loads).
  constexpr int GROUPS = stage_groups(NP, NC);
  constexpr int GSIZE = S::BLOCK / GROUPS;
  constexpr int NPAD = stage_node_pad(NP, NC);
  constexpr int IH = NPAD / GROUPS;
  __shared__ R sQ[EPB * NP * NC];
  __shared__ R sF[ND * NP * EPB * NC];  // [k][j][el][c]; later the rhs [el][i][c]
  __shared__ R sS[NF * NFP * EPB * NC]; // [f*NFP + r][el][c]
  __shared__ R sJ[EPB * ND * ND + EPB]; // invJ then inv_vol
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
  // q_a of this thread's output node (EPB*NP <= 256: one node per thread)
  R qa[NC];
  if (tid < nE * NP) {
    const R *g = A.q_a + (size_t)e0 * NP * NC + (size_t)tid * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) qa[c] = g[c];
  }
  {
    const R *src = A.q_in + (size_t)e0 * NP * NC;
    const int n = nE * NP * NC;
    for (int t = tid; t < n; t += S::BLOCK) sQ[t] = src[t];
    const R *gJ = A.invJ + (size_t)e0 * ND * ND;
    for (int t = tid; t < nE * ND * ND; t += S::BLOCK) sJ[t] = gJ[t];
    for (int t = tid; t < nE; t += S::BLOCK) sJ[EPB * ND * ND + t] = A.inv_vol[e0 + t];
    for (int t = tid; t < NF * NFP; t += S::BLOCK) sMap[t] = A.fmap[t];
  }
  __syncthreads();

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
        sF[((k * NP + j) * EPB + el) * NC + c] = v;
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
    for (int c = 0; c < NC; ++c) sS[((lf * NFP + it.r) * EPB + el) * NC + c] = s * Fs[c];
    if (pin) {
      const int plf = (it.packed >> 16) & 3, rs = A.pslot[(it.packed >> 18) * NFP + it.r];
#pragma unroll
      for (int c = 0; c < NC; ++c) sS[((plf * NFP + rs) * EPB + (pin - 1)) * NC + c] = -(s * Fs[c]);
    }
  }
  __syncthreads();

  // ---- 3a. contractions per (element, component) column ------------------
  // rhs_i = sum_k sum_j D_k[i][j] F_kj - inv_vol sum_fr Lift[fr][i] S_fr,
  // written into sQ (no longer read after the face phase) as [el][i][c].
  {
    const int grp = __builtin_amdgcn_readfirstlane(tid / GSIZE);
    const int i0 = grp * IH;
    for (int p = tid - grp * GSIZE; p < nE * NC; p += GSIZE) {
      const int el = p / NC, c = p - el * NC;
      R acc[IH];
#pragma unroll
      for (int ii = 0; ii < IH; ++ii) acc[ii] = R(0);
#pragma unroll 4
      for (int fr = 0; fr < NF * NFP; ++fr) {
        const R sv = sS[fr * EPB * NC + p];
        const R *L = A.Lt + fr * NPAD + i0;
#pragma unroll
        for (int ii = 0; ii < IH; ++ii) acc[ii] += L[ii] * sv;
      }
      const R niv = -sJ[EPB * ND * ND + el];
#pragma unroll
      for (int ii = 0; ii < IH; ++ii) acc[ii] *= niv;
#pragma unroll 4
      for (int kj = 0; kj < ND * NP; ++kj) {
        const R fv = sF[kj * EPB * NC + p];
        const R *D = A.Dt + kj * NPAD + i0;
#pragma unroll
        for (int ii = 0; ii < IH; ++ii) acc[ii] += D[ii] * fv;
      }
#pragma unroll
      for (int ii = 0; ii < IH; ++ii)
        if (i0 + ii < NP) {
          // the slot holds q_in of this (node, component): fold in b q_in
          R &slot = sQ[(el * NP + i0 + ii) * NC + c];
          slot = A.b * slot + A.cdt * acc[ii];
        }
    }
  }
  __syncthreads();

  // ---- 3b. source, RK combine, limit, store per (element, node) ---------
  if (tid < nE * NP) {
    const int t = tid;
    const size_t g = (size_t)e0 * NP * NC + (size_t)t * NC;
    R src[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) src[c] = R(0);
    if constexpr (Phys::kSource) {
      if (phys.source_active()) {
        R qi[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) qi[c] = A.q_in[g + c];
        phys.source(qi, src);
      }
    }
    R out[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = A.a * qa[c] + sQ[t * NC + c] + A.cdt * src[c];
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
