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
// One block (4 wave64, or 5 for the wide 2D P = 3 tile of the FP32-MFMA
// form, below) per tile of at most EPB elements, which are contiguous in
// device memory:
//   1. each thread starts the loads of its first face item (job record
//      with the face geometry, partner node, and the partner's trace when
//      the partner lies outside the tile); the block stages the q_in tile
//      and the tile's geometry in LDS;
//   2. per face item: the numerical flux (boundary faces: against the
//      physics' ghost state), signed and scaled by the face area, into LDS.
//      A face whose two elements are both in the tile is evaluated once by
//      its side-0 element, which also stores the negated value in the
//      partner's slot; faces to elements outside the tile are evaluated by
//      each side, as the reference evaluates every face from both sides;
//      side 0 is always the left state.  Per (element, node): the nodal
//      flux mapped to reference directions by invJ, into LDS;
//   3. the D and Lift contractions over LDS, then per (element, node) the
//      RK combine, limit_state and the store.
//
// Two arithmetic forms share the kernel:
//   R = double  plain FP64: the contractions run on the VALU in FP64;
//   R = float   FP32-MFMA: FP32 state, fluxes and traces; the volume and
//               LIFT contractions are one [D | Lift] x [F ; S] product per
//               tile on V_MFMA_F32_16X16X4F32 with FP32 accumulation, the
//               surface fluxes pre-scaled by -inv_vol s area; geometry and
//               Jacobians are stored in FP64 (the nodal flux is mapped by
//               invJ rounded to FP32), and the RK combine and the
//               positivity rescue are FP64; the stored state is rounded
//               stochastically.
// Blocks are mapped to tiles so that each of the 8 XCDs walks a contiguous
// range of tiles: face neighbours then mostly hit the XCD's own L2.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "physics.hpp"

namespace mxm {

// Shu-Osher stage q_out = a q_a + b q_in + c dt L(q_in): in every SSPRK3
// stage of src/ssprk3.mojo:41-78 the b-term buffer is the stage input.
template <class R>
struct StageArgs {
  const R *q_in, *q_a;
  R *q_out;
  double a, b, c;         // stage coefficients; c dt multiplies the rhs
  // Instances: a tile belongs to one instance, whose physics parameters
  // (MXM_PRM_MAX doubles), time step and GLM decay the block reads.  One
  // solver is a single instance (tile_inst == nullptr: instance 0 with dt
  // and post_scale below); a batch sets the per-instance tables, and an
  // instance whose inst_active entry is 0 is skipped.
  const int *tile_inst;     // [ntiles] or nullptr
  const double *inst_prm;   // [n_inst][MXM_PRM_MAX]
  const double *inst_dt;    // [n_inst] or nullptr -> dt
  const double *inst_post;  // [n_inst] or nullptr -> post_scale
  const int *inst_active;   // [n_inst] or nullptr (all active)
  int *inst_fail;           // [n_inst] or nullptr: set to 1 when a stored value is not finite
  double *mean;             // [slots][NC] or nullptr: cell means sum_i w_i q_out_i of the stored state
  double *ext;              // [slots][NCL][2] (with mean): smallest and largest stored node value of the
                            // limited components (all NC in 3D, component 0 in 2D)
  const double *w;          // [NP] node weights (with mean)
  // Limiter factors pending on the input buffers (nullptr: none): every
  // value read from q_in (tile and partner traces) or q_a is used as
  // mean + theta (q - mean) where its element's theta < 1, which is the
  // state the limiter would have stored.
  const double *lim_th_in, *lim_mn_in;  // q_in: theta [slots], means [slots][NC]
  const double *lim_th_a, *lim_mn_a;    // q_a
  double dt, post_scale;
  int post_comp;            // component scaled after the combine (GLM psi), -1 none
  const int *tile_count;  // [ntiles] elements in each tile (slots T*EPB ..)
  const double *invJ;     // [slots][ND*ND] rows = reference directions
  const double *inv_vol;  // [slots] 1/(6V) or 1/(2A)
  const double *Dt;       // FP64 form: [ND*NP (k,j)][NPAD (i)] transposed D_ref, zero padded
  const double *Lt;       // FP64 form: [NF*NFP (f,r)][NPAD (i)] transposed Lift_ref, zero padded
  const float *Aop;       // FP32 form: [16*MT (i)][KPAD (k)] = [D_0 .. D_ND-1 | Lift], zero padded
  const int *fmap;        // [NF][NFP] element node of each reference face slot
  const int4 *jobs;       // [ntiles*maxj] {el | lf<<7 | pin<<9 | plf<<16 | pslot<<18
                          //  (-1: padding), partner slot, face id, side | bc<<1}
  const double *jgeo;     // [ntiles*maxj][4] face normal, area
  const unsigned short *ipr;  // [ntiles*maxj*NFP] partner node | partner slot << 8
  int maxj;               // job slots per tile
  int ntiles;
};

constexpr int kPrmMax = 32;  // MXM_PRM_MAX

// Launch configurations of the stage kernel: each is an occupancy target
// (waves per SIMD, e.g. 256-thread blocks per CU) that bounds the register
// allocation; all configurations compute the same bits.  The autotuner
// times them on the problem and keeps the fastest.  FP64 kernels need
// 100-150 VGPRs: 5 waves (<= 102 VGPRs) or 4 (<= 128, which takes the
// 131-VGPR GLM-MHD kernel from 3 waves to 4); the FP32 kernels fit 5, 6
// or 7 waves (<= 72 VGPRs: the 2D trait Euler kernel without spills), and
// the compiler's own choice (1: no bound) can be far off (e.g. 250 VGPRs,
// 1 wave, for the 2D Euler kernels at P = 1).
template <class R>
struct StageCfg;
template <>
struct StageCfg<double> {
  static constexpr int count = 3, dflt = 0;
  static constexpr int waves[3] = {1, 5, 4};
};
template <>
struct StageCfg<float> {
  static constexpr int count = 4, dflt = 1;
  static constexpr int waves[4] = {1, 5, 6, 7};
};
constexpr int kStageCfgMax = 4;

// Elements per tile (one thread per element node): at most 256/NP for NP
// nodes per element, rounded down to whole lattice cells (2 triangles, 6
// tets) when a cell fits, then reduced -- by cells, then to runs of
// elements -- until the block's LDS at FP64 (the operand array -- the q_in
// tile and the face fluxes, later the ND*NP flux planes, rows of
// EPB*NC + 1 values -- and the geometry) fits 64 KB, which only the
// two-fluid state (NC = 17) needs.  Tiles are blocks of cells, so the LDS
// is sized for what a tile can hold; the host builds the tiles with the
// same value.
//
// The wide tile (wide = true; the FP32-MFMA form only, where the solver's
// autotuner times both tile sizes and keeps the faster): when 256 threads
// leave more than 5 % of the lanes without a node and whole cells fill a
// block of 5 to 8 waves exactly -- of the shapes built, 2D P = 3 (NP = 10:
// 24 triangles on 240 of 256 lanes, or 32 on all 320 lanes of 5 waves).
// It pays when all tiles are resident at once: a mesh of 128 x 128
// squares makes 1,024 wide tiles, 4 per CU and 5 waves per SIMD, against
// 1,376 narrow ones, 5 or 6 per CU and 6 waves on the busiest SIMDs.
// When tiles queue for the CUs the narrow tile is faster (an 8-wave tile
// of 50 triangles measured slower still): a CU then holds fewer, larger
// blocks, whose barriers idle more waves at once and whose phases overlap
// less.  The FP64 contraction layout (stage_groups) is sized for 4 waves
// and has no wide tile.
constexpr int stage_lds_bytes(int ND, int NP, int NFP, int NC, int epb) {
  return (epb * NP * NC + (ND + 1) * NFP * (epb * NC + 1) > ND * NP * (epb * NC + 1)
              ? epb * NP * NC + (ND + 1) * NFP * (epb * NC + 1)
              : ND * NP * (epb * NC + 1)) * 8 +
         epb * (ND * ND + 1) * 8 + (ND + 1) * NFP * 4;
}
constexpr int stage_epb(int ND, int NP, int NFP, int NC, bool wide = false) {
  const int cell = ND == 3 ? 6 : 2;
  int e = 256 / NP;
  if (e >= cell) e -= e % cell;
  if (wide && e >= cell && e * NP * 20 < 256 * 19)
    for (int w = 5; w <= 8; ++w) {
      const int t = 64 * w;
      if (t % NP == 0 && (t / NP) % cell == 0 && t / NP <= 126 &&
          stage_lds_bytes(ND, NP, NFP, NC, t / NP) <= 65536)
        return t / NP;
    }
  while (e > 1 && stage_lds_bytes(ND, NP, NFP, NC, e) > 65536) e = e > cell ? e - cell : e - 1;
  return e;
}
// Threads per block: 256, or the whole waves of a wide tile
constexpr int stage_block(int ND, int NP, int NFP, int NC, bool wide = false) {
  return stage_epb(ND, NP, NFP, NC, wide) * NP > 256 ? (stage_epb(ND, NP, NFP, NC, wide) * NP + 63) / 64 * 64 : 256;
}
// whether the shape has a wide tile distinct from the 256-thread one
constexpr bool stage_has_wide(int ND, int NP, int NFP, int NC) {
  return stage_epb(ND, NP, NFP, NC, true) != stage_epb(ND, NP, NFP, NC, false);
}

// FP64 contraction layout: the tile's (element, component) columns span
// ceil(EPB*NC/64) waves; the block's 4 waves form `groups` groups, each
// producing `IH` consecutive output nodes, so the operator entries a wave
// reads are wave-uniform.  The transposed operators are stored with
// groups*IH (>= NP) zero-padded node columns.
constexpr int stage_groups(int EPB, int NC) {
  return (EPB * NC + 63) / 64 >= 4 ? 1 : 4 / ((EPB * NC + 63) / 64);
}
constexpr int stage_node_pad(int EPB, int NP, int NC) {
  return stage_groups(EPB, NC) * ((NP + stage_groups(EPB, NC) - 1) / stage_groups(EPB, NC));
}

// FP32-MFMA operator shape: 16-row tiles of output nodes, K = ND*NP volume
// columns then NF*NFP lift columns, padded to the MFMA depth of 4.
constexpr int mfma_mtiles(int NP) { return (NP + 15) / 16; }
// FP32-MFMA operator columns: the ND*NP derivative columns padded to a
// multiple of 4, then the (ND+1)*NFP lift columns padded to a multiple of 4
constexpr int mfma_kvpad(int ND, int NP) { return (ND * NP + 3) / 4 * 4; }
constexpr int mfma_kpad(int ND, int NP, int NFP) { return mfma_kvpad(ND, NP) + ((ND + 1) * NFP + 3) / 4 * 4; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// Length of the stage kernel's LDS operand array (in values): the q_in
// tile [0, TILE) with the face fluxes after it (NFF rows of NCOLS), and
// later the KV flux-plane rows from 0; the MFMA form also reads its
// padding rows and up to NTC columns of every row.
constexpr int stage_operand_len(bool mfma, int TILE, int NCOLS, int KV, int KVP, int NFF, int KPAD, int NTC) {
  return mfma ? cmax(cmax(TILE + NFF * NCOLS, KVP * NCOLS),
                     cmax(TILE + (KPAD - KVP - 1) * NCOLS + NTC, (KVP - 1) * NCOLS + NTC))
              : cmax(TILE + NFF * NCOLS, KV * NCOLS);
}

template <int ND, int P, int NC = 1, bool WIDE = false>
struct Shape {
  static constexpr int NP = ND == 3 ? (P + 1) * (P + 2) * (P + 3) / 6 : (P + 1) * (P + 2) / 2;
  static constexpr int NFP = ND == 3 ? (P + 1) * (P + 2) / 2 : P + 1;
  static constexpr int NF = ND + 1;
  static constexpr int BLOCK = stage_block(ND, NP, NFP, NC, WIDE);
  static constexpr int EPB = stage_epb(ND, NP, NFP, NC, WIDE);
  static constexpr bool kHasWide = stage_has_wide(ND, NP, NFP, NC);
};
// the stage kernel's shape for the state type R (the wide tile: FP32 only)
template <class Phys, int ND, int P, class R, bool WIDE>
using StageShape = Shape<ND, P, Phys::NC, WIDE && std::is_same<R, float>::value>;
template <class Phys, int ND, int P, class R, bool WIDE>
constexpr int kStageBlock = StageShape<Phys, ND, P, R, WIDE>::BLOCK;

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
  int packed, r, side, bc, pn, rs, pl;
  R n[3];
  double area;
  R qp[NC];  // partner trace when the partner is outside the tile
};

// A value of the state with its element's pending limiter factor applied:
// the expression of the limiter's own apply, so the bits are those the
// limiter would have stored.
template <class R>
__device__ __forceinline__ R lim_apply(R v, double th, double bm) {
  return th < 1.0 ? R(bm + th * (double(v) - bm)) : v;
}

// Every record of face item t of a tile is addressed from the tile index,
// so all of them are in flight together (load_item_records); the partner
// trace of a face to another tile is the only load that waits for one of
// them (load_item_trace).
template <int NFP, class R, int NC>
__device__ __forceinline__ void load_item_records(const StageArgs<R> &A, int tile, int t, int e0,
                                                  FaceItem<R, NC> &it) {
  const int jb = t / NFP;
  it.r = t - jb * NFP;
  const size_t j = (size_t)tile * A.maxj + jb;
  const int4 jr = A.jobs[j];
  const double2 g0 = reinterpret_cast<const double2 *>(A.jgeo)[2 * j];
  const double2 g1 = reinterpret_cast<const double2 *>(A.jgeo)[2 * j + 1];
  const unsigned pr = A.ipr[(size_t)tile * A.maxj * NFP + t];
  it.packed = jr.x;
  it.side = jr.w & 1;
  it.bc = (jr.w >> 1) & 7;
  it.n[0] = R(g0.x), it.n[1] = R(g0.y), it.n[2] = R(g1.x);
  it.area = g1.y;
  it.pn = pr & 255;
  it.rs = pr >> 8;
  it.pl = jr.y - e0;
}

// A partner slot in [0, EPB) is in the tile: partners are real elements,
// which occupy the tile's first slots.
template <int NP, int EPB, bool LIM, class R, int NC>
__device__ __forceinline__ void load_item_trace(const StageArgs<R> &A, int e0, FaceItem<R, NC> &it) {
  const bool need = it.packed >= 0 && !(it.side == 0 && it.bc != 0);
  if (need && !(it.pl >= 0 && it.pl < EPB)) {
    const size_t p = (size_t)(e0 + it.pl);
    const R *g = A.q_in + (p * NP + it.pn) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) it.qp[c] = g[c];
    if constexpr (LIM) {
      const double th = A.lim_th_in[p];
#pragma unroll
      for (int c = 0; c < NC; ++c) it.qp[c] = lim_apply(it.qp[c], th, A.lim_mn_in[p * NC + c]);
    }
  }
}

typedef float mxm_float4 __attribute__((ext_vector_type(4)));

// FP64 -> FP32 with stochastic rounding, for the FP32-MFMA form's stored
// state.  Round-to-nearest storage loses every update smaller than half a
// float ulp and biases the rest, so over thousands of steps on a nearly
// steady background the state drifts from the FP64 solution; rounding up
// or down at random, with probabilities that make the expected stored
// value equal x, leaves only an unbiased random walk.  The 29 mantissa
// bits a float drops are cleared (the float of what is left is exact) and
// one float ulp is added away from zero when a 32-bit draw is below those
// bits scaled to 32 bits (lo << 3): probability (dropped bits) / 2^29.
// The draw is the bit reversal of x's low word xor `salt` (one value per
// node of an element), so its high bits are x's lowest, noisiest bits: a
// pure function of x and its node, so results do not depend on the launch
// configuration or on the other instances of a batch.  Infinities and the
// hardware's NaN have a zero low word and convert unchanged; any other NaN
// stays non-finite.
__device__ __forceinline__ float sr_f32(double x, unsigned salt) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)b;
  const float t = float(__longlong_as_double((long long)(b & ~0x1FFFFFFFull)));
  const unsigned up = __builtin_bitreverse32(lo ^ salt) < (lo << 3) ? 1u : 0u;
  return __uint_as_float(__float_as_uint(t) + up);
}

// W: occupancy target of the launch configuration, in waves per SIMD
// (amdgpu_waves_per_eu); 1 leaves the register budget to the compiler.
// WIDE: the wide tile (stage_epb).
template <class Phys, int ND, int P, class R, int W, bool LIM, bool WIDE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kStageBlock<Phys, ND, P, R, WIDE>),
                          amdgpu_waves_per_eu(W))) void
dg_stage_kernel(StageArgs<R> A) {
  using S = StageShape<Phys, ND, P, R, WIDE>;
  constexpr bool kMFMA = std::is_same<R, float>::value;
  static_assert(kMFMA || S::BLOCK == 256, "the FP64 contraction layout is sized for 4 waves");
  constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, EPB = S::EPB, NC = Phys::NC;
  constexpr int TILE = EPB * NP * NC;  // values of one element tile
  constexpr int NCOL = EPB * NC;       // (element, component) columns of a tile
  // Rows of (element, component) columns are NCOLS apart: one word of
  // padding, so that the node threads' row-strided writes of the flux
  // planes spread over the LDS banks.
  constexpr int NCOLS = NCOL + 1;
  // The contraction's right operands: NFF = NF*NFP face-flux rows, then
  // KV = ND*NP flux-plane rows in the same array (stage_operand_len).
  // FP32-MFMA: the products run over 4-deep steps of operator columns and
  // NT 16-column tiles; the array is long enough for every such read, and
  // the padding rows (face fluxes NFF..KPAD-KVP-1, planes KV..KVP-1) hold
  // zeros (their operator columns are zero; a column beyond NCOL only
  // feeds an output column that is never stored).
  constexpr int NFF = NF * NFP, KV = ND * NP;
  constexpr int KVP = kMFMA ? mfma_kvpad(ND, NP) : KV;
  constexpr int KPAD = kMFMA ? mfma_kpad(ND, NP, NFP) : KV + NFF;
  constexpr int MT = mfma_mtiles(NP);
  constexpr int NT = (NCOL + 15) / 16;
  // 2D kernels without the limiter, FP32 form with 7 or more components
  // (GLM-MHD) and FP64 form with at most 5: q_a is loaded with the tile
  // into LDS past the operand array and read by the node threads in 2b,
  // instead of waiting in NC registers through the face phase, the
  // register peak.  FP32 GLM-MHD then fits 6 waves per SIMD without spills
  // (its 384^2 stage 2 % faster) and the FP64 Euler kernels 5 (96 VGPRs,
  // 31.8 KB LDS: the 512^2 vortex stage 4 % faster, C2 0.5 %).  The FP32
  // Euler kernels lost 1 % (their registers were not the limit), FP64
  // GLM-MHD and Maxwell would become LDS-bound, and the 3D kernels' LDS
  // would bound their blocks per CU.
  constexpr bool kQaLds = !LIM && ND == 2 && (kMFMA ? NC >= 7 : NC <= 5);
  constexpr int NB0 = stage_operand_len(kMFMA, TILE, NCOLS, KV, KVP, NFF, KPAD, NT * 16);
  constexpr int QA0 = NB0;  // past every operand read and the zeroed padding
  constexpr int NB = NB0 + (kQaLds ? TILE : 0);
  // FP32-MFMA: the operator [D | Lift] staged in LDS when it is at most
  // 16 KB (2D, 3D P <= 3), else read from global memory (L2) in the loop
  constexpr int NA = kMFMA ? MT * 16 * KPAD : 0;
  constexpr bool kStageA = kMFMA && NA * 4 <= 16384;
  // LDS row stride of the staged operator: 4 mod 8 words, so the 16 rows
  // x 4 columns an A operand read touches fall in 64 distinct banks (a
  // stride of 32, as 2D P = 3 has, put all 16 rows in one bank group)
  constexpr int KAS = KPAD % 8 == 4 ? KPAD : KPAD + 4;
  // sW = sB: first the q_in tile as [el][j][c] (TILE contiguous values);
  // after the lift product the contravariant nodal fluxes as [k][j][el][c]
  // (row kj at kj*NCOLS); after the volume product the result as
  // [el][i][c].  sS = sB + TILE: face fluxes [f*NFP + r][el][c].
  __shared__ R sB[NB];
  __shared__ float sA[kStageA ? MT * 16 * KAS : 1];
  __shared__ double sJ[EPB * ND * ND + EPB];  // invJ then inv_vol
  __shared__ int sMap[NF * NFP];
  // pending limiter factors of the tile's elements, then their means (q_in)
  constexpr int NL = EPB * (NC + 1);
  __shared__ double sLim[LIM ? NL : 1];
  R *const sW = sB;
  R *const sS = sB + TILE;

  const int tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int inst = A.tile_inst ? A.tile_inst[tile] : 0;
  if (A.inst_active && !A.inst_active[inst]) return;
  const int e0 = tile * EPB;
  const Phys phys = Phys::from(A.inst_prm + (size_t)inst * kPrmMax);
  const double cdt = A.c * (A.inst_dt ? A.inst_dt[inst] : A.dt);
  const double post = A.inst_post ? A.inst_post[inst] : A.post_scale;
  const int nitems = A.maxj * NFP;  // padding items have packed < 0

  // ---- 1. loads --------------------------------------------------------
  // First everything addressed from the tile index -- the q_in tile, the
  // geometry, the first face item's records and q_a of the own node --
  // then the partner trace that waits for the records; the tile and the
  // geometry go to LDS last.  Every tile owns EPB element slots in every
  // per-slot array (padding slots hold zeros), so these loads cover the
  // whole tile and need not wait for the tile's element count.
  constexpr int QPT = (TILE + S::BLOCK - 1) / S::BLOCK;  // tile values per thread
  constexpr int NJ = EPB * ND * ND, NG = NJ + EPB;      // invJ, then inv_vol
  constexpr int GPT = (NG + S::BLOCK - 1) / S::BLOCK;
  R tq[QPT];
  {
    const R *src = A.q_in + (size_t)e0 * NP * NC;
#pragma unroll
    for (int k = 0; k < QPT; ++k)
      if (k * S::BLOCK + tid < TILE) tq[k] = src[k * S::BLOCK + tid];
  }
  const bool qa_is_in = A.q_a == A.q_in;
  R tqa[kQaLds ? QPT : 1];
  if constexpr (kQaLds) {
    if (!qa_is_in) {
      const R *src = A.q_a + (size_t)e0 * NP * NC;
#pragma unroll
      for (int k = 0; k < QPT; ++k)
        if (k * S::BLOCK + tid < TILE) tqa[k] = src[k * S::BLOCK + tid];
    }
  }
  double tg[GPT];
#pragma unroll
  for (int k = 0; k < GPT; ++k) {
    const int t = k * S::BLOCK + tid;
    if (t < NJ)
      tg[k] = A.invJ[(size_t)e0 * ND * ND + t];
    else if (t < NG)
      tg[k] = A.inv_vol[e0 + t - NJ];
  }
  const int fm = tid < NF * NFP ? A.fmap[tid] : 0;
  constexpr int APT = kStageA ? (NA + S::BLOCK - 1) / S::BLOCK : 0;
  float ta[APT > 0 ? APT : 1];
  if constexpr (kStageA) {
#pragma unroll
    for (int k = 0; k < APT; ++k)
      if (k * S::BLOCK + tid < NA) ta[k] = A.Aop[k * S::BLOCK + tid];
  }
  constexpr int LPT = (NL + S::BLOCK - 1) / S::BLOCK;
  double tl[LIM ? LPT : 1];
  if constexpr (LIM) {
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
      const int t = k * S::BLOCK + tid;
      if (t < EPB)
        tl[k] = A.lim_th_in[e0 + t];
      else if (t < NL)
        tl[k] = A.lim_mn_in[(size_t)e0 * NC + t - EPB];
    }
  }
  FaceItem<R, NC> it;
  if (tid < nitems) load_item_records<NFP>(A, tile, tid, e0, it);
  // q_a of the own node (from LDS in 2b when kQaLds), in the state
  // precision until 2b; when q_a is the stage input (SSPRK3 stage 1) it is
  // read from the LDS copy of the tile
  R qa_in[NC];
  if (!kQaLds && tid < EPB * NP && !qa_is_in) {
    const R *g = A.q_a + (size_t)e0 * NP * NC + (size_t)tid * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) qa_in[c] = g[c];
  }
  const int nE = A.tile_count[tile];
  // thread tid owns output node (el, i) = (tid / NP, tid % NP): EPB*NP <= BLOCK
  const bool node = tid < nE * NP;
  if (LIM && !qa_is_in && tid < EPB * NP) {
    const size_t e = (size_t)e0 + tid / NP;
    const double th = A.lim_th_a[e];
#pragma unroll
    for (int c = 0; c < NC; ++c) qa_in[c] = lim_apply(qa_in[c], th, A.lim_mn_a[e * NC + c]);
  }
  if (tid < nitems) load_item_trace<NP, EPB, LIM>(A, e0, it);
#pragma unroll
  for (int k = 0; k < QPT; ++k)
    if (k * S::BLOCK + tid < TILE) sW[k * S::BLOCK + tid] = tq[k];
#pragma unroll
  for (int k = 0; k < GPT; ++k) {
    const int t = k * S::BLOCK + tid;
    if (t < NG) sJ[t] = tg[k];
  }
  if (tid < NF * NFP) sMap[tid] = fm;
  if constexpr (kQaLds) {
    if (!qa_is_in) {
#pragma unroll
      for (int k = 0; k < QPT; ++k)
        if (k * S::BLOCK + tid < TILE) sB[QA0 + k * S::BLOCK + tid] = tqa[k];
    }
  }
  if constexpr (kStageA) {
#pragma unroll
    for (int k = 0; k < APT; ++k) {
      const int t = k * S::BLOCK + tid, row = t / KPAD;
      if (t < NA) sA[row * KAS + t - row * KPAD] = ta[k];
    }
  }
  if constexpr (kMFMA)
    for (int t = TILE + NFF * NCOLS + tid; t < NB0; t += S::BLOCK) sB[t] = R(0);
  if constexpr (LIM) {
#pragma unroll
    for (int k = 0; k < LPT; ++k)
      if (k * S::BLOCK + tid < NL) sLim[k * S::BLOCK + tid] = tl[k];
  }
  __syncthreads();
  // the q_in tile's pending limiter factors, applied in place to the
  // values this thread stored
  if constexpr (LIM) {
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      const int t = k * S::BLOCK + tid;
      if (t < TILE) {
        const int el = t / (NP * NC), c = t % NC;
        sW[t] = lim_apply(tq[k], sLim[el], sLim[EPB + el * NC + c]);
      }
    }
    __syncthreads();
  }
  const R *sQ = sW;  // the q_in tile, [el][j][c], until the flux planes are written

  // ---- 2a. face fluxes ----------------------------------------------------
  for (int t = tid; t < nitems; t += S::BLOCK) {
    if (t != tid) {
      load_item_records<NFP>(A, tile, t, e0, it);
      load_item_trace<NP, EPB, LIM>(A, e0, it);
    }
    if (it.packed < 0) continue;
    const int el = it.packed & 127, lf = (it.packed >> 7) & 3, pin = (it.packed >> 9) & 127;
    R qm[NC], qo[NC], ql[NC], qr[NC];
    const int mine = (el * NP + sMap[lf * NFP + it.r]) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) qm[c] = sQ[mine + c];
    if (it.pl >= 0 && it.pl < EPB) {
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
    phys.template numerical_flux<ND>(ql, qr, it.n, Fs);
    const double sa = it.side == 0 ? it.area : -it.area;
    // FP64 form: s area F*, lifted and scaled by -inv_vol in the
    // contraction; FP32 form: -inv_vol s area F*, ready for the product
    const R own = kMFMA ? R(-sJ[EPB * ND * ND + el] * sa) : R(sa);
#pragma unroll
    for (int c = 0; c < NC; ++c) sS[(lf * NFP + it.r) * NCOLS + el * NC + c] = own * Fs[c];
    if (pin) {
      const int plf = (it.packed >> 16) & 3, rs = it.rs;
      const R oth = kMFMA ? R(sJ[EPB * ND * ND + pin - 1] * sa) : R(-sa);
#pragma unroll
      for (int c = 0; c < NC; ++c) sS[(plf * NFP + rs) * NCOLS + (pin - 1) * NC + c] = oth * Fs[c];
    }
  }

  // ---- 2b. nodal flux of the own node, and the RK terms of q_in ----------
  // qa becomes a q_a + b q_in + c dt S(q_in) (FP64); the flux, mapped to
  // reference directions by invJ in FP64, waits in registers until every
  // read of the q_in tile is done.
  R Fk[ND * NC];
  double qa[NC];
  if (node) {
    const int el = tid / NP;
    const R *q = &sQ[tid * NC];
    R F[ND * NC];
    phys.template flux<ND>(q, F);
    const double *iJ = &sJ[el * ND * ND];
    if constexpr (kMFMA) {
      // FP32 form: the mapping in FP32 (invJ rounded once per node), as
      // the volume product that takes its result is
      float jf[ND * ND];
#pragma unroll
      for (int t = 0; t < ND * ND; ++t) jf[t] = float(iJ[t]);
#pragma unroll
      for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float v = 0.f;
#pragma unroll
          for (int d = 0; d < ND; ++d) v += jf[k * ND + d] * F[d * NC + c];
          Fk[k * NC + c] = v;
        }
    } else {
#pragma unroll
      for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double v = 0.0;
#pragma unroll
          for (int d = 0; d < ND; ++d) v += iJ[k * ND + d] * double(F[d * NC + c]);
          Fk[k * NC + c] = R(v);
        }
    }
    R src[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) src[c] = R(0);
    if constexpr (Phys::kSource) {
      if (phys.source_active()) phys.source(q, src);
    }
    if constexpr (kQaLds) {
      if (!qa_is_in) {
#pragma unroll
        for (int c = 0; c < NC; ++c) qa_in[c] = sB[QA0 + tid * NC + c];
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      qa[c] = A.a * double(qa_is_in ? q[c] : qa_in[c]) + A.b * double(q[c]) + cdt * double(src[c]);
  }
  __syncthreads();

  // ---- 3. contractions ----------------------------------------------------
  // rhs_i = sum_k sum_j D_k[i][j] F_kj - inv_vol sum_fr Lift[fr][i] S_fr in
  // two products: first the lift over the face fluxes, then -- once the
  // face fluxes and the q_in tile are no longer read -- the flux planes are
  // written over them and the volume product accumulates onto the same
  // registers.  The result goes to [el][i][c] at the start of the array.
  //
  // FP32-MFMA: Out[i][col] = sum_k Aop[i][k] B[k][col] with the operator
  // columns [D (KVP) | Lift (KPAD - KVP)]; V_MFMA_F32_16X16X4F32 operands,
  // lane l = 16 g + r: A[r][k0 + g], B[k0 + g][col0 + r]; result rows
  // 4 g + v (v < 4), column col0 + r.
  constexpr int NTW = (NT + S::BLOCK / 64 - 1) / (S::BLOCK / 64);  // column tiles per wave
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l = tid & 63, lr = l & 15, lg = l >> 4;
  mxm_float4 macc[kMFMA ? MT : 1][kMFMA ? NTW : 1];
  // FP64 on the VALU: the (element, component) columns are the lanes;
  // each wave group produces IH consecutive output nodes, so the operator
  // entries are wave-uniform (scalar loads)
  constexpr int GROUPS = stage_groups(EPB, NC);
  constexpr int GSIZE = S::BLOCK / GROUPS;
  constexpr int NPAD = stage_node_pad(EPB, NP, NC);
  constexpr int IH = NPAD / GROUPS;
  constexpr int ITER = (NCOL + GSIZE - 1) / GSIZE;  // columns per thread, at most
  const int grp = __builtin_amdgcn_readfirstlane(tid / GSIZE);
  const int i0 = grp * IH;
  const int lane = tid - grp * GSIZE;
  double dacc[kMFMA ? 1 : ITER][kMFMA ? 1 : IH];

  // one 4-deep step of the MFMA product over operator columns k0.. and
  // operand rows kb.. of Bb
  auto kstep = [&](const R *Bb, int k0, int kb) {
    float a[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      if constexpr (kStageA)
        a[m] = sA[(m * 16 + lr) * KAS + k0 + lg];
      else
        a[m] = A.Aop[(m * 16 + lr) * KPAD + k0 + lg];
    }
#pragma unroll
    for (int w = 0; w < NTW; ++w) {
      const int nt = wave * NTW + w;
      if (nt < NT) {
        const float bv = float(Bb[(kb + lg) * NCOLS + nt * 16 + lr]);
#pragma unroll
        for (int m = 0; m < MT; ++m) macc[m][w] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], bv, macc[m][w], 0, 0, 0);
      }
    }
  };

  // ---- 3a. lift product over the face fluxes -----------------------------
  if constexpr (kMFMA) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int w = 0; w < NTW; ++w) macc[m][w] = mxm_float4{0.f, 0.f, 0.f, 0.f};
    if constexpr (kStageA) {
#pragma unroll
      for (int k0 = KVP; k0 < KPAD; k0 += 4) kstep(sS, k0, k0 - KVP);
    } else {
#pragma unroll 4
      for (int k0 = KVP; k0 < KPAD; k0 += 4) kstep(sS, k0, k0 - KVP);
    }
  } else {
#pragma unroll
    for (int m = 0; m < ITER; ++m) {
      const int p = lane + m * GSIZE;
#pragma unroll
      for (int ii = 0; ii < IH; ++ii) dacc[m][ii] = 0.0;
      if (p < nE * NC) {
        const int el = p / NC;
#pragma unroll 4
        for (int fr = 0; fr < NFF; ++fr) {
          const double sv = sS[fr * NCOLS + p];
          const double *L = A.Lt + fr * NPAD + i0;
#pragma unroll
          for (int ii = 0; ii < IH; ++ii) dacc[m][ii] += L[ii] * sv;
        }
        const double niv = -sJ[EPB * ND * ND + el];
#pragma unroll
        for (int ii = 0; ii < IH; ++ii) dacc[m][ii] *= niv;
      }
    }
  }
  __syncthreads();

  // ---- 3b. flux planes over the face fluxes and the q_in tile -------------
  if (node) {
    const int el = tid / NP, j = tid - el * NP;
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) sW[(k * NP + j) * NCOLS + el * NC + c] = Fk[k * NC + c];
  }
  if constexpr (kMFMA && KVP > KV)
    for (int t = KV * NCOLS + tid; t < KVP * NCOLS; t += S::BLOCK) sB[t] = R(0);
  __syncthreads();

  // ---- 3c. volume product, result as [el][i][c] ---------------------------
  if constexpr (kMFMA) {
    if constexpr (kStageA) {
#pragma unroll
      for (int k0 = 0; k0 < KVP; k0 += 4) kstep(sW, k0, k0);
    } else {
#pragma unroll 4
      for (int k0 = 0; k0 < KVP; k0 += 4) kstep(sW, k0, k0);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NTW; ++w) {
      const int col = (wave * NTW + w) * 16 + lr;
      if (wave * NTW + w < NT && col < nE * NC) {
        const int el = col / NC, c = col - el * NC;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = m * 16 + 4 * lg + v;
            if (i < NP) sW[(el * NP + i) * NC + c] = R(macc[m][w][v]);
          }
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < ITER; ++m) {
      const int p = lane + m * GSIZE;
      if (p < nE * NC) {
#pragma unroll 4
        for (int kj = 0; kj < KV; ++kj) {
          const double fv = sW[kj * NCOLS + p];
          const double *D = A.Dt + kj * NPAD + i0;
#pragma unroll
          for (int ii = 0; ii < IH; ++ii) dacc[m][ii] += D[ii] * fv;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < ITER; ++m) {
      const int p = lane + m * GSIZE;
      if (p < nE * NC) {
        const int el = p / NC, c = p - el * NC;
#pragma unroll
        for (int ii = 0; ii < IH; ++ii)
          if (i0 + ii < NP) sW[(el * NP + i0 + ii) * NC + c] = R(dacc[m][ii]);
      }
    }
  }
  __syncthreads();

  // ---- 4. RK combine (FP64), limit per (element, node), tile store -------
  // Each node thread writes its finished values over its own entries of the
  // [el][i][c] result (the tile's device order); after a barrier the block
  // stores the tile's real elements as one contiguous run, every lane
  // consecutive, so each cache line is written whole at once (per-node
  // stores at a stride of NC leave partly written lines in L2 that are
  // written back more than once when many blocks are resident).
  if (node) {
    double out[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = qa[c] + cdt * double(sW[tid * NC + c]);
    if constexpr (Phys::kLimit) phys.limit(out);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c == A.post_comp) out[c] *= post;
    const unsigned salt = unsigned(tid - (tid / NP) * NP) * 0x85EBCA6Bu;  // node in its element
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (kMFMA)
        sW[tid * NC + c] = sr_f32(out[c], salt);
      else
        sW[tid * NC + c] = R(out[c]);
    }
  }
  __syncthreads();
  {
    R *g = A.q_out + (size_t)e0 * NP * NC;
    const int nv = nE * NP * NC;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      const int t = k * S::BLOCK + tid;
      if (t < nv) {
        const R v = sW[t];
        g[t] = v;
        bad = bad || !__builtin_isfinite(v);
      }
    }
    if (bad && A.inst_fail) atomicOr(&A.inst_fail[inst], 1);
  }
  // ---- 5. cell means of the stored state (limiter) -----------------------
  // sum_i w_i q_i per (element, component) in node order, FP64, from the
  // stored values: what a separate pass over q_out would compute
  if constexpr (LIM) {
    for (int p = tid; p < nE * NC; p += S::BLOCK) {
      const int el = p / NC, c = p - el * NC;
      double sm = 0.0;
#pragma unroll
      for (int n = 0; n < NP; ++n) sm += A.w[n] * double(sW[(el * NP + n) * NC + c]);
      A.mean[(size_t)(e0 + el) * NC + c] = sm;
      constexpr int NCL = ND == 3 ? NC : 1;  // the components the limiter bounds
      if (c < NCL) {
        double lo = double(sW[el * NP * NC + c]), hi = lo;
#pragma unroll
        for (int n = 1; n < NP; ++n) {
          const double v = double(sW[(el * NP + n) * NC + c]);
          lo = v < lo ? v : lo;
          hi = v > hi ? v : hi;
        }
        A.ext[((size_t)(e0 + el) * NCL + c) * 2] = lo;
        A.ext[((size_t)(e0 + el) * NCL + c) * 2 + 1] = hi;
      }
    }
  }
}

// A shape without a distinct wide tile launches its one kernel for WIDE.
template <class Phys, int ND, int P, class R, int W, bool LIM, bool WIDE>
hipError_t launch_stage(const void *args, hipStream_t s) {
  constexpr bool kWide = WIDE && std::is_same<R, float>::value && Shape<ND, P, Phys::NC>::kHasWide;
  if constexpr (WIDE && !kWide) {
    return launch_stage<Phys, ND, P, R, W, LIM, false>(args, s);
  } else {
    using S = StageShape<Phys, ND, P, R, WIDE>;
    const StageArgs<R> &a = *static_cast<const StageArgs<R> *>(args);
    const int nb = a.ntiles;
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL((dg_stage_kernel<Phys, ND, P, R, W, LIM, WIDE>), dim3(nb), dim3(S::BLOCK), 0, s, a);
    return hipGetLastError();
  }
}

// args points to StageArgs<double> (FP64 form) or StageArgs<float>
// (FP32-MFMA form); the tables are in stage_families.hpp.
using StageLaunch = hipError_t (*)(const void *args, hipStream_t);

}  // namespace mxm
