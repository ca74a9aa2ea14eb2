// batch.hip -- the batch engine: many independent DG problems advanced
// together, one stage launch per kernel group and RK stage.
//
// An instance is one problem (mesh, order, physics parameters, time step,
// step count, initial state).  Instances that use the same stage kernel
// (dimension, P, physics, flux variant, limiter) form a group whose tiles
// share one device arena: the stage kernel's block reads its tile's
// instance slot and from it the physics parameters and time step
// (StageArgs::tile_inst / inst_*), so one launch advances every instance
// of the group by one stage.  A device step counter per instance slot
// turns the instance off (inst_active = 0) after its last step, or as soon
// as one of its stored values is not finite (inst_fail, set by the stage
// kernel), so instances with different step counts share the launches and
// finished or failed ones cost a block exit.  One SSPRK3 step of a group
// (three stages, limiter, counter bump) is a HIP graph, replayed.
//
// Every group has its own HIP stream: groups (different physics, orders,
// limiter settings) run concurrently on the device.  A new instance is
// placed into a free tile range of its group's arena -- a range a removed
// instance left, or the arena's unused tail -- by asynchronous copies on
// the group's stream, with no synchronisation.  Only when no range fits
// (or the instance needs more face-job slots per tile than the arena has)
// is the group compacted: its stream is drained, a new arena is built for
// its placed and waiting instances (longest remaining run first, so the
// grid shrinks as instances finish), and the old one is released.  The other
// groups keep running meanwhile.
//
// Results do not depend on the mix: a tile never mixes instances, and
// every element's arithmetic only reads its own instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/mxm_dg.h"
#include "cfl.hpp"
#include "diag.hpp"
#include "limiter.hpp"
#include "runtime.hpp"

namespace mxm {
namespace {

// One SSPRK3 step done for every active slot; a slot whose state went
// non-finite is switched off instead.
__global__ void bump_steps_kernel(int *steps, const int *nsteps, int *active, const int *fail, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && active[i]) {
    if (fail[i]) {
      active[i] = 0;
      return;
    }
    const int s = steps[i] + 1;
    steps[i] = s;
    if (s >= nsteps[i]) active[i] = 0;
  }
}

// dst[0..n) = v (tile ranges handed back to the dead slot, switched-off
// slots); a kernel rather than a memset so that it is ordinary stream work
__global__ void fill_int_kernel(int *dst, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

struct BInst {
  int id = 0, dim = 0, P = 0, phys = 0, variant = 0, NC = 0, K = 0, NP = 0;
  bool limiter = false;
  double eps = 0.1;
  double prm[MXM_PRM_MAX] = {0};
  double dt = 0.0;
  int nsteps = 0;
  int steps_done = 0;      // host mirror of the device counter
  int steps_enqueued = 0;  // steps launched for it so far (the host's own count)
  bool failed = false, reported = false;
  bool placed = false;
  int base_tile = 0, slot = 0;
  std::vector<double> q0;  // initial state, reference order, until placed
  ElemTables t;
  int ntiles() const { return t.til.ntiles; }
  bool live() const { return !failed && steps_enqueued < nsteps; }
};

using GroupKey = std::tuple<int, int, int, int, int, long long>;  // dim, P, phys, variant, limiter, eps bits

// The device arena of a kernel group, sized for FreeSpace::cap tiles and
// cap_slots instance slots: state trio, per-slot and per-job tables,
// per-instance scalars.
struct Arena {
  DevBuf<double> qd[3];
  DevBuf<float> qf[3];
  // limiter (groups with it): cell means [3][slots*NC] and factors
  // [3][slots] of the stage buffers, node extrema [slots*NCL][2] (the
  // limited components: all NC in 3D, component 0 in 2D)
  DevBuf<double> lmean, ltheta, lext;
  DevBuf<double> invJ, inv_vol, jgeo, iprm, idt, ipost;
  DevBuf<int> conn;  // limiter: partner slot per face
  DevBuf<int4> jobs;
  DevBuf<unsigned short> ipr;
  DevBuf<int> tile_count, tile_inst, iactive, isteps, insteps, ifail;
  void swap(Arena &o) noexcept {
    for (int k = 0; k < 3; ++k) qd[k].swap(o.qd[k]), qf[k].swap(o.qf[k]);
    lmean.swap(o.lmean), ltheta.swap(o.ltheta), lext.swap(o.lext), invJ.swap(o.invJ), inv_vol.swap(o.inv_vol), jgeo.swap(o.jgeo), iprm.swap(o.iprm);
    idt.swap(o.idt), ipost.swap(o.ipost), conn.swap(o.conn), jobs.swap(o.jobs), ipr.swap(o.ipr);
    tile_count.swap(o.tile_count), tile_inst.swap(o.tile_inst), iactive.swap(o.iactive), isteps.swap(o.isteps);
    insteps.swap(o.insteps), ifail.swap(o.ifail);
  }
  size_t bytes() const {
    size_t b = 0;
    for (int k = 0; k < 3; ++k) b += qd[k].bytes() + qf[k].bytes();
    return b + lmean.bytes() + ltheta.bytes() + lext.bytes() + invJ.bytes() + inv_vol.bytes() + jgeo.bytes() + iprm.bytes() + idt.bytes() +
           ipost.bytes() + conn.bytes() + jobs.bytes() + ipr.bytes() + tile_count.bytes() + tile_inst.bytes() +
           iactive.bytes() + isteps.bytes() + insteps.bytes() + ifail.bytes();
  }
};

// Pinned host memory for asynchronous uploads; busy until its event (on
// the stream that reads it) completes.
struct Staging {
  char *p = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool busy = false;
  ~Staging() {
    if (ev) (void)hipEventDestroy(ev);
    if (p) (void)hipHostFree(p);
  }
};

// Free tile ranges of an arena: holes inside [0, top) and the unused tail
// [top, cap).
struct FreeSpace {
  int top = 0, cap = 0;
  std::map<int, int> holes;  // start -> length
  // a free range of nt tiles: a hole, else the tail; -1 if none
  int find(int nt) const {
    for (auto &h : holes)
      if (h.second >= nt) return h.first;
    return top + nt <= cap ? top : -1;
  }
  void take(int start, int nt) {
    if (start == top) {
      top += nt;
      return;
    }
    auto it = holes.find(start);
    const int len = it->second;
    holes.erase(it);
    if (len > nt) holes.emplace(start + nt, len - nt);
  }
  void give(int start, int nt) {
    auto it = holes.emplace(start, nt).first;
    auto nx = std::next(it);
    if (nx != holes.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      holes.erase(nx);
    }
    if (it != holes.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        holes.erase(it);
        it = pv;
      }
    }
    if (it->first + it->second == top) {
      top = it->first;
      holes.erase(it);
    }
  }
};

struct BGroup {
  int dim = 0, P = 0, phys = 0, variant = 0, NC = 0, NP = 0, NFP = 0, NF = 0, epb = 0, prec = 0;
  int cfg = 0;  // launch configuration (StageCfg), autotuned per grid size
  bool limiter = false;
  double eps = 0.1;
  StageLaunch launch = nullptr;
  GroupKey key{};
  hipStream_t stream = nullptr;
  Arena ar;
  int cap_slots = 0, maxj = 1;
  int compactions = 0;          // arenas built so far
  FreeSpace fs;                 // tiles of the arena not held by placed instances
  int next_slot = 1;            // slots [1, next_slot) are taken; slot 0 is the dead slot
  std::vector<int> placed;      // instances in the arena
  std::vector<int> waiting;     // added, not yet placed
  int grid_tiles = 0;           // tiles the captured graphs cover
  DevBuf<double> Dt, Lt, w;
  DevBuf<double> Dn, dpart, dout;  // diagnostics
  DevBuf<float> Aop;
  DevBuf<int> fmap;
  // replayed HIP graphs of kGraphSteps steps ([0]) and of one step ([1]);
  // graphs replaced while their launches may still run wait in `retired`
  // until the group's stream has finished that work
  static constexpr int kGraphSteps = 8;
  StepGraph sg[2];
  std::vector<StepGraph> retired;
  // pinned mirrors of the device step counters and failure flags, copied
  // asynchronously at the end of every step() call (two buffers, so the
  // copy of one call and the host's read of the previous one never meet)
  int *h_cnt = nullptr;  // [2][2 * h_cap]: steps, then failure flags
  int h_cap = 0;
  int h_n[2] = {0, 0};   // slots each buffer holds
  hipEvent_t cev[2] = {nullptr, nullptr};
  long long calls = 0;   // step() calls that recorded counters since the last compaction
  long long seen = 0;    // counters of calls [0, seen) have been read

  ~BGroup() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &e : sg) e.reset();
    for (auto &r : retired) r.reset();
    for (auto &e : cev)
      if (e) (void)hipEventDestroy(e);
    if (h_cnt) (void)hipHostFree(h_cnt);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void retire_graphs() {
    for (auto &e : sg)
      if (e.exec) {
        retired.push_back(e);
        e = StepGraph{};
      }
  }
  size_t bytes() const {
    return ar.bytes() + Dn.bytes() + dpart.bytes() + dout.bytes() + Dt.bytes() + Lt.bytes() + w.bytes() +
           Aop.bytes() + fmap.bytes();
  }
};

// Device bytes one instance adds to its group's arena (exact: the tables
// its placement writes, with the instance's own job slots per tile).
size_t instance_bytes(const ElemTables &t, int NP, int NC, int NFP, int dim, int prec, bool limiter) {
  const size_t slots = (size_t)t.til.slots(), tiles = (size_t)t.til.ntiles, js = tiles * t.fj.maxj;
  const size_t word = prec == MXM_PREC_FP64 ? sizeof(double) : sizeof(float);
  size_t b = 3 * slots * NP * NC * word;                            // state trio
  b += slots * (dim * dim + 1) * sizeof(double);                    // invJ, inv_vol
  b += slots * (dim + 1) * sizeof(int);                             // limiter neighbours
  b += js * (sizeof(int4) + 4 * sizeof(double) + NFP * sizeof(unsigned short));  // jobs, geometry, partner nodes
  b += tiles * 2 * sizeof(int);                                     // tile_count, tile_inst
  b += (MXM_PRM_MAX + 2) * sizeof(double) + 4 * sizeof(int);        // per-instance scalars
  if (limiter) b += slots * (3 * NC + 2 * (dim == 3 ? NC : 1) + 3) * sizeof(double);  // means, factors, extrema
  return b;
}

// Host images of one instance's tables at its place in an arena (base
// tile bt, slot sl, maxj job slots per tile), in the arena's layouts.
struct InstImage {
  std::vector<double> ij, iv, geo, prm;
  std::vector<int> cn;
  std::vector<int4> jb;
  std::vector<unsigned short> pr;
  std::vector<int> tc, ti;
  double dt = 0.0, post = 1.0;
  int active = 0, steps = 0, nsteps = 0;
};

InstImage inst_image(const BGroup &g, const BInst &in, int bt, int sl, int maxj) {
  InstImage m;
  const ElemTables &t = in.t;
  const int epb = g.epb, dd = g.dim * g.dim, NF = g.NF, NFP = g.NFP;
  const int lt = t.til.ntiles, lmj = t.fj.maxj, bs = bt * epb;
  const size_t slots = (size_t)lt * epb;
  m.ij.assign(t.invJ.begin(), t.invJ.begin() + slots * dd);
  m.iv.assign(t.inv_vol.begin(), t.inv_vol.begin() + slots);
  m.cn.resize(slots * NF);
  for (size_t s = 0; s < slots * NF; ++s) m.cn[s] = t.conn[s] + bs;
  m.jb.resize((size_t)lt * maxj);
  m.geo.assign((size_t)lt * maxj * 4, 0.0);
  m.pr.assign((size_t)lt * maxj * NFP, 0);
  m.tc.resize(lt);
  m.ti.assign(lt, sl);
  for (int T = 0; T < lt; ++T) {
    m.tc[T] = t.til.count[T];
    for (int j = 0; j < maxj; ++j) {
      const size_t dj = (size_t)T * maxj + j;
      if (j < lmj) {
        const size_t sj = (size_t)T * lmj + j;
        m.jb[dj] = int4{t.fj.jobs[sj * 4], t.fj.jobs[sj * 4 + 1] + bs, t.fj.jobs[sj * 4 + 2], t.fj.jobs[sj * 4 + 3]};
        for (int c = 0; c < 4; ++c) m.geo[dj * 4 + c] = t.fj.geo[sj * 4 + c];
        for (int r = 0; r < NFP; ++r) m.pr[dj * NFP + r] = t.fj.ipr[sj * NFP + r];
      } else {
        m.jb[dj] = int4{-1, (bt + T) * epb, 0, 0};  // padding job: in-bounds partner, skipped
      }
    }
  }
  m.prm.assign(in.prm, in.prm + MXM_PRM_MAX);
  m.dt = in.dt;
  m.post = post_factor(in.phys, in.prm, in.dt);
  m.steps = in.steps_done;
  m.nsteps = in.nsteps;
  m.active = (!in.failed && in.steps_done < in.nsteps) ? 1 : 0;
  return m;
}

// the instance's state in device order, zero-filled padding slots
template <class R>
std::vector<R> device_state(const BGroup &g, const BInst &in) {
  const size_t node = (size_t)g.NP * g.NC, nv = (size_t)in.ntiles() * g.epb * node;
  std::vector<R> d(nv, R(0));
  for (int e = 0; e < in.K; ++e)
    for (size_t c = 0; c < node; ++c) d[(size_t)in.t.til.dev_of[e] * node + c] = R(in.q0[(size_t)e * node + c]);
  return d;
}

struct Batch {
  int prec = MXM_PREC_FP64;
  hipStream_t cstream = nullptr;  // graph capture only (never runs work)
  std::map<GroupKey, std::unique_ptr<BGroup>> groups;
  std::map<int, std::unique_ptr<BInst>> insts;
  std::map<std::pair<int, int>, RefElem> refs;
  std::vector<std::unique_ptr<Staging>> staging;
  int next_id = 0;
  // autotuned launch configuration per (group, grid-size octave)
  bool tune = true;
  // steps as replayed HIP graphs (default) or as direct stream launches of
  // the same kernels
  bool graphs = true;
  // extra capacity of an arena built by a compaction after the first, as a
  // fraction of its tiles and slots (0: exact)
  double headroom = 0.0;
  std::map<std::pair<GroupKey, int>, int> tuned;

  ~Batch() {
    groups.clear();  // drains each group's stream
    staging.clear();
    if (cstream) (void)hipStreamDestroy(cstream);
  }

  const RefElem &ref(int dim, int P) {
    auto k = std::make_pair(dim, P);
    auto it = refs.find(k);
    if (it == refs.end()) it = refs.emplace(k, build_refelem(dim, P)).first;
    return it->second;
  }

  BInst &inst(int id) {
    auto it = insts.find(id);
    if (it == insts.end()) throw ArgError("unknown instance id");
    return *it->second;
  }

  BGroup &group_of_key(const BInst &in) {
    long long eb = 0;
    std::memcpy(&eb, &in.eps, sizeof(eb));
    const GroupKey key{in.dim, in.P, in.phys, in.variant, in.limiter ? 1 : 0, in.limiter ? eb : 0};
    auto it = groups.find(key);
    if (it != groups.end()) return *it->second;
    auto g = std::make_unique<BGroup>();
    const RefElem &re = ref(in.dim, in.P);
    g->dim = in.dim, g->P = in.P, g->phys = in.phys, g->variant = in.variant, g->NC = in.NC, g->NP = re.NP,
    g->NFP = re.NFP, g->NF = re.NF;
    // 256-thread tiles: an arena's tiles queue for the CUs, where the wide
    // tile is slower (stage_epb)
    g->epb = stage_epb(in.dim, re.NP, re.NFP, in.NC);
    g->prec = prec;
    g->limiter = in.limiter, g->eps = in.eps;
    g->key = key;
    g->cfg = stage_cfg_default(prec);
    g->launch = find_stage(in.dim, in.phys, in.variant, in.P, prec, g->cfg, g->limiter);
    if (!g->launch) throw ArgError("no stage kernel for this (physics, flux, P)");
    OpTables o = build_operators(re, in.dim, in.NC, prec);
    if (prec == MXM_PREC_FP64) {
      g->Dt.upload(o.Dt.data(), o.Dt.size());
      g->Lt.upload(o.Lt.data(), o.Lt.size());
    } else {
      g->Aop.upload(o.Aop.data(), o.Aop.size());
    }
    g->w.upload(re.w.data(), re.w.size());
    g->Dn.upload(re.Dn.data(), re.Dn.size());
    g->fmap.upload(re.fmap.data(), re.fmap.size());
    MXM_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (auto &e : g->cev) MXM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return *groups.emplace(key, std::move(g)).first->second;
  }

  BGroup &group_with(int id) {
    const BInst &in = inst(id);
    long long eb = 0;
    std::memcpy(&eb, &in.eps, sizeof(eb));
    const GroupKey key{in.dim, in.P, in.phys, in.variant, in.limiter ? 1 : 0, in.limiter ? eb : 0};
    auto it = groups.find(key);
    if (it == groups.end()) throw ArgError("instance has no group");
    return *it->second;
  }

  int add(int dim, const void *desc, int P, int phys, const double *prm, double dt, int nsteps, int limiter,
          double eps, const double *q0, long long n) {
    if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
    const int nc = phys_nc(phys);
    if (nc < 0) throw ArgError("unknown physics id");
    if (dim == 3 && phys > MXM_PHYS_EULER && phys < MXM_PHYS_SW) throw ArgError("2D module physics on a 3D mesh");
    if (dim == 2 && phys >= MXM_PHYS_SW) throw ArgError("3D physics on a 2D mesh");
    if (!(dt > 0.0) || nsteps < 0) throw ArgError("bad time step or step count");
    if (limiter && nc > kLimMaxNC) throw ArgError("limiter supports NC <= " + std::to_string(kLimMaxNC));
    auto in = std::make_unique<BInst>();
    in->dim = dim, in->P = P, in->phys = phys, in->NC = nc;
    if (prm) std::memcpy(in->prm, prm, sizeof(in->prm));
    in->variant = phys_variant(phys, in->prm);
    in->dt = dt, in->nsteps = nsteps, in->limiter = limiter != 0, in->eps = eps;
    const RefElem &re = ref(dim, P);
    HostMesh m;
    if (dim == 3) {
      auto d = static_cast<const mxm_mesh3d_desc *>(desc);
      check_bc(d->bc, 6);
      m = build_mesh_3d(d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->bc, P, re);
    } else {
      auto d = static_cast<const mxm_mesh2d_desc *>(desc);
      check_bc(d->bc, 4);
      m = build_mesh_2d(d->nx, d->ny, d->lx, d->ly, d->bc, P, re);
    }
    in->K = m.K, in->NP = m.NP;
    if (!q0 || n != (long long)m.K * m.NP * nc) throw ArgError("initial state length must be K*NP*NC");
    in->q0.assign(q0, q0 + n);
    in->t = build_elem_tables(m, re, nc, false);
    in->id = next_id++;
    BGroup &g = group_of_key(*in);
    g.waiting.push_back(in->id);
    const int id = in->id;
    insts.emplace(id, std::move(in));
    return id;
  }

  // ---- device step counters ---------------------------------------------
  void ensure_counter_mirror(BGroup &g) {
    if (g.h_cap >= g.cap_slots && g.h_cnt) return;
    if (g.h_cnt) (void)hipHostFree(g.h_cnt);
    g.h_cnt = nullptr;
    g.h_cap = std::max(g.cap_slots, 16);
    MXM_HIP(hipHostMalloc(reinterpret_cast<void **>(&g.h_cnt), (size_t)2 * 2 * g.h_cap * sizeof(int)));
  }

  // host mirrors from counter buffer b (valid: its copy has completed)
  void absorb_counters(BGroup &g, int b) {
    const int *st = g.h_cnt + (size_t)b * 2 * g.h_cap, *fl = st + g.h_cap;
    for (int id : g.placed) {
      BInst &in = inst(id);
      if (in.slot >= g.h_n[b]) continue;  // placed after that copy
      in.steps_done = std::max(in.steps_done, st[in.slot]);
      if (fl[in.slot]) in.failed = true;
    }
  }

  // synchronous read of the counters of every placed instance (the
  // group's stream must be idle)
  void pull_counters(BGroup &g) {
    if (g.placed.empty() || !g.ar.isteps.p) return;
    std::vector<int> st(g.next_slot), fl(g.next_slot);
    MXM_HIP(hipMemcpy(st.data(), g.ar.isteps.p, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    MXM_HIP(hipMemcpy(fl.data(), g.ar.ifail.p, fl.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int id : g.placed) {
      BInst &in = inst(id);
      in.steps_done = st[in.slot];
      if (fl[in.slot]) in.failed = true;
    }
  }

  void drain(BGroup &g) {
    MXM_HIP(hipStreamSynchronize(g.stream));
    for (auto &r : g.retired) r.reset();
    g.retired.clear();
  }

  // writes the limited values of slots [s0, s0 + n) of the group's state q
  // and clears their pending factors (stream order: after the steps
  // enqueued so far)
  void materialize(BGroup &g, int s0, int n) {
    const size_t S = g.ar.ltheta.n / 3;
    if (n <= 0 || (size_t)s0 + n > S) return;
    const size_t vo = (size_t)s0 * g.NP * g.NC;
    if (g.prec == MXM_PREC_FP64)
      MXM_HIP(launch_bj_apply<double>(g.ar.qd[0].p + vo, g.ar.ltheta.p + s0, g.ar.lmean.p + (size_t)s0 * g.NC, n,
                                      g.NP, g.NC, g.stream));
    else
      MXM_HIP(launch_bj_apply<float>(g.ar.qf[0].p + vo, g.ar.ltheta.p + s0, g.ar.lmean.p + (size_t)s0 * g.NC, n,
                                     g.NP, g.NC, g.stream));
  }

  // ---- staging ----------------------------------------------------------
  Staging &stage_buf(size_t bytes) {
    for (auto &s : staging)
      if (s->busy && hipEventQuery(s->ev) == hipSuccess) s->busy = false;
    for (auto &s : staging)
      if (!s->busy && s->cap >= bytes) return *s;
    auto s = std::make_unique<Staging>();
    s->cap = std::max<size_t>(bytes, 1 << 20);
    MXM_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->p), s->cap));
    MXM_HIP(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming));
    staging.push_back(std::move(s));
    return *staging.back();
  }

  // ---- placement --------------------------------------------------------
  // asynchronous upload of one instance at (bt, sl) into the arena
  template <class R>
  void upload_instance(BGroup &g, DevBuf<R> (&q)[3], BInst &in, int bt, int sl) {
    const InstImage m = inst_image(g, in, bt, sl, g.maxj);
    const std::vector<R> qs = device_state<R>(g, in);
    const int dd = g.dim * g.dim, NF = g.NF;
    const size_t bs = (size_t)bt * g.epb;
    struct Seg {
      void *dst;
      const void *src;
      size_t bytes;
    };
    const Seg segs[] = {
        {q[0].p + bs * g.NP * g.NC, qs.data(), qs.size() * sizeof(R)},
        {g.ar.invJ.p + bs * dd, m.ij.data(), m.ij.size() * sizeof(double)},
        {g.ar.inv_vol.p + bs, m.iv.data(), m.iv.size() * sizeof(double)},
        {g.ar.conn.p + bs * NF, m.cn.data(), m.cn.size() * sizeof(int)},
        {g.ar.jobs.p + (size_t)bt * g.maxj, m.jb.data(), m.jb.size() * sizeof(int4)},
        {g.ar.jgeo.p + (size_t)bt * g.maxj * 4, m.geo.data(), m.geo.size() * sizeof(double)},
        {g.ar.ipr.p + (size_t)bt * g.maxj * g.NFP, m.pr.data(), m.pr.size() * sizeof(unsigned short)},
        {g.ar.tile_count.p + bt, m.tc.data(), m.tc.size() * sizeof(int)},
        {g.ar.tile_inst.p + bt, m.ti.data(), m.ti.size() * sizeof(int)},
        {g.ar.iprm.p + (size_t)sl * MXM_PRM_MAX, m.prm.data(), MXM_PRM_MAX * sizeof(double)},
        {g.ar.idt.p + sl, &m.dt, sizeof(double)},
        {g.ar.ipost.p + sl, &m.post, sizeof(double)},
        {g.ar.isteps.p + sl, &m.steps, sizeof(int)},
        {g.ar.insteps.p + sl, &m.nsteps, sizeof(int)},
        {g.ar.iactive.p + sl, &m.active, sizeof(int)},
    };
    size_t total = 0;
    for (const Seg &s : segs) total += (s.bytes + 15) / 16 * 16;
    Staging &st = stage_buf(total);
    size_t off = 0;
    for (const Seg &s : segs) {
      std::memcpy(st.p + off, s.src, s.bytes);
      MXM_HIP(hipMemcpyAsync(s.dst, st.p + off, s.bytes, hipMemcpyHostToDevice, g.stream));
      off += (s.bytes + 15) / 16 * 16;
    }
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, g.stream, g.ar.ifail.p + sl, 0, 1);
    MXM_HIP(hipGetLastError());
    if (g.limiter) {  // nothing pending on the new state (the range may hold a removed instance's factors)
      const size_t S = g.ar.ltheta.n / 3;
      for (int b = 0; b < 3; ++b)
        MXM_HIP(launch_fill(g.ar.ltheta.p + b * S + bs, (long long)in.ntiles() * g.epb, 1.0, g.stream));
    }
    MXM_HIP(hipEventRecord(st.ev, g.stream));
    st.busy = true;
  }

  // places the group's waiting instances into free ranges (asynchronous);
  // compacts the group when they do not all fit
  void place(BGroup &g) {
    if (g.waiting.empty()) return;
    std::vector<int> order = g.waiting;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst(a).nsteps > inst(b).nsteps; });
    // fits without a compaction?
    bool fits = g.fs.cap > 0 && g.next_slot + (int)order.size() <= g.cap_slots;
    if (fits) {
      FreeSpace trial = g.fs;
      for (int id : order) {
        const BInst &in = inst(id);
        const int r = in.t.fj.maxj <= g.maxj ? trial.find(in.ntiles()) : -1;
        if (r < 0) {
          fits = false;
          break;
        }
        trial.take(r, in.ntiles());
      }
    }
    if (!fits) {
      compact(g);
      return;
    }
    for (int id : order) {
      BInst &in = inst(id);
      const int bt = g.fs.find(in.ntiles());
      g.fs.take(bt, in.ntiles());
      const int sl = g.next_slot++;
      if (g.prec == MXM_PREC_FP64)
        upload_instance(g, g.ar.qd, in, bt, sl);
      else
        upload_instance(g, g.ar.qf, in, bt, sl);
      in.base_tile = bt, in.slot = sl, in.placed = true;
      in.steps_enqueued = in.steps_done;
      in.q0.clear();
      in.q0.shrink_to_fit();
      g.placed.push_back(id);
    }
    g.waiting.clear();
  }

  // Rebuilds the group's arena for its live and waiting instances, sized
  // exactly for them.  Everything is built in a new arena first; only when
  // every allocation and copy has succeeded are the arena and the
  // instances' places replaced, so an error leaves the group as it was
  // (its waiting instances still waiting).
  void compact(BGroup &g) {
    drain(g);
    if (g.limiter && g.ar.ltheta.n) {  // the states move with their limiter factors applied
      materialize(g, 0, (int)(g.ar.ltheta.n / 3));
      drain(g);
    }
    pull_counters(g);
    // every placed instance stays (finished ones remain readable until
    // they are removed), longest remaining run first: finished and failed
    // instances end up in the tail, which the grid leaves out
    std::vector<int> order = g.placed;
    order.insert(order.end(), g.waiting.begin(), g.waiting.end());
    auto remaining = [&](int id) {
      const BInst &x = inst(id);
      return x.failed ? 0 : std::max(0, x.nsteps - x.steps_done);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return remaining(a) > remaining(b); });
    int tiles = 0, maxj = 1;
    for (int id : order) {
      tiles += inst(id).ntiles();
      maxj = std::max(maxj, inst(id).t.fj.maxj);
    }
    // a group compacted before has grown: give it headroom (a fraction of
    // its tiles and slots) so that later instances are placed without one
    const bool grow = g.compactions > 0 && headroom > 0.0;
    const int ni = (int)order.size();
    const int cap_slots = ni + 1 + (grow ? std::max(8, (int)(ni * headroom)) : 0);
    const int cap_tiles = std::max(1, tiles + (grow ? (int)(tiles * headroom) : 0));
    const int epb = g.epb, dd = g.dim * g.dim, NF = g.NF, NFP = g.NFP;
    const size_t slots = (size_t)cap_tiles * epb, tile_vals = (size_t)epb * g.NP * g.NC;
    Arena na;
    auto state_trio = [&](auto &nq, auto &oq) {
      using R = typename std::remove_reference<decltype(*nq[0].p)>::type;
      for (auto &b : nq) {
        b.alloc(slots * g.NP * g.NC);
        MXM_HIP(hipMemset(b.p, 0, b.bytes()));
      }
      int bt = 0;
      for (int id : order) {
        const BInst &in = inst(id);
        R *dst = nq[0].p + (size_t)bt * tile_vals;
        const size_t nv = (size_t)in.ntiles() * tile_vals;
        if (in.placed) {
          MXM_HIP(hipMemcpy(dst, oq[0].p + (size_t)in.base_tile * tile_vals, nv * sizeof(R),
                            hipMemcpyDeviceToDevice));
        } else {
          const std::vector<R> d = device_state<R>(g, in);
          MXM_HIP(hipMemcpy(dst, d.data(), nv * sizeof(R), hipMemcpyHostToDevice));
        }
        bt += in.ntiles();
      }
    };
    if (g.prec == MXM_PREC_FP64)
      state_trio(na.qd, g.ar.qd);
    else
      state_trio(na.qf, g.ar.qf);
    std::vector<double> ij(slots * dd, 0.0), iv(slots, 0.0), geo((size_t)cap_tiles * maxj * 4, 0.0);
    std::vector<int> cn(slots * NF);
    std::vector<int4> jb((size_t)cap_tiles * maxj);
    std::vector<unsigned short> pr((size_t)cap_tiles * maxj * NFP, 0);
    std::vector<int> tc(cap_tiles, 0), ti(cap_tiles, 0);
    std::vector<double> iprm((size_t)cap_slots * MXM_PRM_MAX, 0.0), idt(cap_slots, 0.0), ipost(cap_slots, 1.0);
    std::vector<int> iact(cap_slots, 0), ist(cap_slots, 0), inst_n(cap_slots, 0), ifl(cap_slots, 0);
    for (int s = 0; s < (int)slots; ++s)
      for (int lf = 0; lf < NF; ++lf) cn[(size_t)s * NF + lf] = s;
    for (int T = 0; T < cap_tiles; ++T)
      for (int j = 0; j < maxj; ++j) jb[(size_t)T * maxj + j] = int4{-1, T * epb, 0, 0};
    std::vector<int> new_base(ni);
    int bt = 0;
    for (int k = 0; k < ni; ++k) {
      const BInst &in = inst(order[k]);
      const int sl = k + 1;
      const InstImage m = inst_image(g, in, bt, sl, maxj);
      const size_t bs = (size_t)bt * epb;
      std::copy(m.ij.begin(), m.ij.end(), ij.begin() + bs * dd);
      std::copy(m.iv.begin(), m.iv.end(), iv.begin() + bs);
      std::copy(m.cn.begin(), m.cn.end(), cn.begin() + bs * NF);
      std::copy(m.jb.begin(), m.jb.end(), jb.begin() + (size_t)bt * maxj);
      std::copy(m.geo.begin(), m.geo.end(), geo.begin() + (size_t)bt * maxj * 4);
      std::copy(m.pr.begin(), m.pr.end(), pr.begin() + (size_t)bt * maxj * NFP);
      std::copy(m.tc.begin(), m.tc.end(), tc.begin() + bt);
      std::copy(m.ti.begin(), m.ti.end(), ti.begin() + bt);
      std::copy(m.prm.begin(), m.prm.end(), iprm.begin() + (size_t)sl * MXM_PRM_MAX);
      idt[sl] = m.dt, ipost[sl] = m.post, iact[sl] = m.active, ist[sl] = m.steps, inst_n[sl] = m.nsteps;
      new_base[k] = bt;
      bt += in.ntiles();
    }
    na.invJ.upload(ij.data(), ij.size());
    na.inv_vol.upload(iv.data(), iv.size());
    na.conn.upload(cn.data(), cn.size());
    na.jobs.upload(jb.data(), jb.size());
    na.jgeo.upload(geo.data(), geo.size());
    na.ipr.upload(pr.data(), pr.size());
    na.tile_count.upload(tc.data(), tc.size());
    na.tile_inst.upload(ti.data(), ti.size());
    na.iprm.upload(iprm.data(), iprm.size());
    na.idt.upload(idt.data(), idt.size());
    na.ipost.upload(ipost.data(), ipost.size());
    na.iactive.upload(iact.data(), iact.size());
    na.isteps.upload(ist.data(), ist.size());
    na.insteps.upload(inst_n.data(), inst_n.size());
    na.ifail.upload(ifl.data(), ifl.size());
    if (g.limiter) {
      na.lmean.alloc(3 * slots * g.NC);
      na.lext.alloc(2 * slots * (g.dim == 3 ? g.NC : 1));
      na.ltheta.alloc(3 * slots);
      MXM_HIP(launch_fill(na.ltheta.p, (long long)na.ltheta.n, 1.0, g.stream));
    }
    // ---- commit (no more allocations or copies that can fail) ----
    g.ar.swap(na);
    for (auto &e : g.sg) e.reset();
    g.grid_tiles = 0;
    g.cap_slots = cap_slots, g.maxj = maxj;
    g.fs = FreeSpace{};
    g.fs.top = tiles;
    g.fs.cap = cap_tiles;
    g.next_slot = ni + 1;
    ++g.compactions;
    for (int k = 0; k < ni; ++k) {
      BInst &in = inst(order[k]);
      in.base_tile = new_base[k], in.slot = k + 1, in.placed = true;
      in.steps_enqueued = in.steps_done;
      if (!in.q0.empty()) {
        in.q0.clear();
        in.q0.shrink_to_fit();
      }
    }
    g.placed = order;
    g.waiting.clear();
    g.calls = g.seen = 0;
    ensure_counter_mirror(g);
  }

  // ---- stepping ---------------------------------------------------------
  // one past the last tile of the instances still to run at least one of
  // the next steps
  int live_top(const BGroup &g) {
    int t = 0;
    for (int id : g.placed) {
      const BInst &in = inst(id);
      if (in.live()) t = std::max(t, in.base_tile + in.ntiles());
    }
    return t;
  }

  // graphs over the first `tiles` tiles (none when 0); the old ones are
  // retired until the work launched from them is done
  void recapture(BGroup &g, int tiles) {
    g.retire_graphs();
    g.grid_tiles = tiles;
    if (tiles > 0) {
      if (tune) choose_config(g);
      if (graphs) capture(g);
    }
  }

  // the launch configuration for the group's grid size: from the cache of
  // this (group, tile-count octave), else timed on stage 2 (q1 and q ->
  // q2, after one stage 1 q -> q1; the next step rewrites q1 and q2 before
  // reading them)
  void choose_config(BGroup &g) {
    int oct = 0;
    while ((1 << oct) < g.grid_tiles) ++oct;
    const auto key = std::make_pair(g.key, oct);
    auto it = tuned.find(key);
    if (it == tuned.end()) {
      int c;
      if (g.prec == MXM_PREC_FP64) {
        const StageArgs<double> a1 = stage_args(g, g.ar.qd, 0), a = stage_args(g, g.ar.qd, 1);
        MXM_HIP(g.launch(&a1, g.stream));
        c = autotune_stage(g.dim, g.phys, g.variant, g.P, g.prec, &a, g.stream, 10, nullptr, g.limiter);
      } else {
        const StageArgs<float> a1 = stage_args(g, g.ar.qf, 0), a = stage_args(g, g.ar.qf, 1);
        MXM_HIP(g.launch(&a1, g.stream));
        c = autotune_stage(g.dim, g.phys, g.variant, g.P, g.prec, &a, g.stream, 10, nullptr, g.limiter);
      }
      it = tuned.emplace(key, c).first;
    }
    g.cfg = it->second;
    g.launch = find_stage(g.dim, g.phys, g.variant, g.P, g.prec, g.cfg, g.limiter);
  }

  template <class R>
  StageArgs<R> stage_args(const BGroup &g, const DevBuf<R> (&q)[3], int s) const {
    StageArgs<R> a{};
    const StagePlan pl = stage_plan(s);
    const Arena &ar = g.ar;
    a.q_in = q[pl.in].p, a.q_a = q[pl.qa].p, a.q_out = q[pl.out].p;
    a.a = pl.a, a.b = pl.b, a.c = pl.c;
    a.tile_inst = ar.tile_inst.p, a.inst_prm = ar.iprm.p, a.inst_dt = ar.idt.p, a.inst_post = ar.ipost.p;
    a.inst_active = ar.iactive.p;
    a.inst_fail = ar.ifail.p;
    if (g.limiter) {  // as in the solver: outputs' means and extrema, inputs' pending factors
      const size_t S = ar.ltheta.n / 3, kn = S * g.NC;
      a.mean = ar.lmean.p + pl.out * kn, a.ext = ar.lext.p;
      a.lim_th_in = ar.ltheta.p + pl.in * S, a.lim_mn_in = ar.lmean.p + pl.in * kn;
      a.lim_th_a = ar.ltheta.p + pl.qa * S, a.lim_mn_a = ar.lmean.p + pl.qa * kn;
    }
    a.w = g.w.p;
    a.dt = 0.0, a.post_scale = 1.0;
    // GLM decay per instance (factor 1 where alpha_d = 0)
    a.post_comp = s == 2 && g.phys == MXM_PHYS_MHDGLM2D ? 6 : -1;
    a.tile_count = ar.tile_count.p;
    a.invJ = ar.invJ.p, a.inv_vol = ar.inv_vol.p;
    a.Dt = g.Dt.p, a.Lt = g.Lt.p, a.Aop = g.Aop.p, a.fmap = g.fmap.p;
    a.jobs = ar.jobs.p, a.jgeo = ar.jgeo.p, a.ipr = ar.ipr.p, a.maxj = g.maxj, a.ntiles = g.grid_tiles;
    return a;
  }

  template <class R>
  void launch_step(BGroup &g, DevBuf<R> (&q)[3], hipStream_t s) {
    for (int st = 0; st < 3; ++st) {
      const StageArgs<R> a = stage_args(g, q, st);
      MXM_HIP(g.launch(&a, s));
      if (g.limiter)
        MXM_HIP(launch_bj_theta(g.ar.ltheta.p + (g.ar.ltheta.n / 3) * stage_plan(st).out, a.mean, g.ar.lext.p,
                                g.ar.conn.p, g.grid_tiles * g.epb, g.NC, g.NF, g.dim, g.eps, s, g.ar.tile_count.p,
                                g.epb, g.ar.tile_inst.p, g.ar.iactive.p));
    }
    // over every slot of the arena (later placements need no recapture;
    // slots not in use are inactive)
    hipLaunchKernelGGL(bump_steps_kernel, dim3((g.cap_slots + 255) / 256), dim3(256), 0, s, g.ar.isteps.p,
                       g.ar.insteps.p, g.ar.iactive.p, g.ar.ifail.p, g.cap_slots);
    MXM_HIP(hipGetLastError());
  }

  // n SSPRK3 steps of the group as one HIP graph (kernel nodes only),
  // captured on the capture stream; the graph is kept beside its executable
  void capture_steps(BGroup &g, int n, StepGraph &out) {
    MXM_HIP(hipStreamBeginCapture(cstream, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < n; ++k) {
        if (g.prec == MXM_PREC_FP64)
          launch_step(g, g.ar.qd, cstream);
        else
          launch_step(g, g.ar.qf, cstream);
      }
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(cstream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    MXM_HIP(hipStreamEndCapture(cstream, &out.graph));
    MXM_HIP(hipGraphInstantiate(&out.exec, out.graph, nullptr, nullptr, 0));
    // the executable goes to the device now, not at its first launch, on
    // the stream that launches it
    MXM_HIP(hipGraphUpload(out.exec, g.stream));
  }
  void capture(BGroup &g) {
    capture_steps(g, BGroup::kGraphSteps, g.sg[0]);
    capture_steps(g, 1, g.sg[1]);
  }

  // k steps of the group from graph `which` (k = its step count), or the
  // same launches made directly on the group's stream
  void launch_graph(BGroup &g, int which, int k) {
    const int need = live_top(g);
    if (need == 0) return;
    if (need > g.grid_tiles || need <= g.grid_tiles - g.grid_tiles / 8 || (graphs && !g.sg[which].exec))
      recapture(g, need);
    if (graphs) {
      MXM_HIP(hipGraphLaunch(g.sg[which].exec, g.stream));
    } else {
      for (int j = 0; j < k; ++j) {
        if (g.prec == MXM_PREC_FP64)
          launch_step(g, g.ar.qd, g.stream);
        else
          launch_step(g, g.ar.qf, g.stream);
      }
    }
    for (int id : g.placed) {
      BInst &in = inst(id);
      if (in.live()) in.steps_enqueued = std::min(in.nsteps, in.steps_enqueued + k);
    }
  }

  // the counters of this call, copied into the pinned mirror behind the
  // steps just launched
  void record_counters(BGroup &g) {
    if (g.placed.empty()) return;
    ensure_counter_mirror(g);
    const int b = (int)(g.calls & 1);
    int *dst = g.h_cnt + (size_t)b * 2 * g.h_cap;
    MXM_HIP(hipMemcpyAsync(dst, g.ar.isteps.p, (size_t)g.next_slot * sizeof(int), hipMemcpyDeviceToHost, g.stream));
    MXM_HIP(hipMemcpyAsync(dst + g.h_cap, g.ar.ifail.p, (size_t)g.next_slot * sizeof(int), hipMemcpyDeviceToHost,
                           g.stream));
    MXM_HIP(hipEventRecord(g.cev[b], g.stream));
    g.h_n[b] = g.next_slot;
    ++g.calls;
  }

  // releases graphs retired before a call whose work is known done
  void reap(BGroup &g) {
    if (g.retired.empty()) return;
    if (hipStreamQuery(g.stream) != hipSuccess) return;
    for (auto &r : g.retired) r.reset();
    g.retired.clear();
  }

  void step(int n) {
    for (auto &kv : groups) {
      BGroup &g = *kv.second;
      reap(g);
      place(g);
      int m = n;
      for (; m >= BGroup::kGraphSteps; m -= BGroup::kGraphSteps) launch_graph(g, 0, BGroup::kGraphSteps);
      for (; m > 0; --m) launch_graph(g, 1, 1);
      record_counters(g);
    }
  }

  // Instances newly finished (or failed).  lag = 0 waits for all launched
  // work; lag = 1 waits only for the work of step() calls before the most
  // recent one (the most recent keeps the device busy meanwhile).
  std::vector<int> poll(int lag) {
    std::vector<int> out;
    for (auto &kv : groups) {
      BGroup &g = *kv.second;
      if (lag <= 0 || g.calls < 2) {
        drain(g);
        pull_counters(g);
      } else {
        const long long c = g.calls - 2;  // the call before the most recent
        if (c >= g.seen) {
          const int b = (int)(c & 1);
          MXM_HIP(hipEventSynchronize(g.cev[b]));
          absorb_counters(g, b);
          g.seen = c + 1;
        }
      }
    }
    for (auto &kv : insts) {
      BInst &in = *kv.second;
      if (!in.reported && in.placed && (in.failed || in.steps_done >= in.nsteps)) {
        out.push_back(in.id);
        in.reported = true;
      }
    }
    return out;
  }

  void sync_all() {
    for (auto &kv : groups) {
      drain(*kv.second);
      pull_counters(*kv.second);
    }
  }

  template <class R>
  void read_state(BGroup &g, DevBuf<R> (&q)[3], BInst &in, double *h) {
    const size_t node = (size_t)g.NP * g.NC, tile_vals = (size_t)g.epb * node;
    const size_t nv = (size_t)in.ntiles() * tile_vals;
    std::vector<R> d(nv);
    MXM_HIP(hipMemcpy(d.data(), q[0].p + (size_t)in.base_tile * tile_vals, nv * sizeof(R), hipMemcpyDeviceToHost));
    for (int e = 0; e < in.K; ++e)
      for (size_t c = 0; c < node; ++c) h[(size_t)e * node + c] = double(d[(size_t)in.t.til.dev_of[e] * node + c]);
  }

  // the group of a placed instance, drained (placement first if needed)
  BGroup &settled(int id) {
    BGroup &g = group_with(id);
    if (!inst(id).placed) place(g);
    const BInst &in = inst(id);
    if (g.limiter && in.placed) materialize(g, in.base_tile * g.epb, in.ntiles() * g.epb);
    drain(g);
    return g;
  }

  void get_state(int id, double *h, long long n) {
    BInst &in = inst(id);
    if (!h || n != (long long)in.K * in.NP * in.NC) throw ArgError("state length must be K*NP*NC");
    if (!in.placed && !in.q0.empty()) {
      std::memcpy(h, in.q0.data(), n * sizeof(double));
      return;
    }
    BGroup &g = settled(id);
    if (g.prec == MXM_PREC_FP64)
      read_state(g, g.ar.qd, in, h);
    else
      read_state(g, g.ar.qf, in, h);
  }

  // the largest signal speed over the nodes of one instance
  double max_wave_speed(int id) {
    BInst &in = inst(id);
    BGroup &g = settled(id);
    const int nt = in.ntiles();
    if (g.dpart.n < (size_t)nt) g.dpart.alloc(nt);
    if (g.dout.n < 1) g.dout.alloc(1);
    auto fill = [&](auto &a) {
      a.prm = g.ar.iprm.p, a.tile_inst = g.ar.tile_inst.p, a.tile_count = g.ar.tile_count.p;
      a.epb = g.epb, a.NP = g.NP, a.tile0 = in.base_tile, a.ntiles = nt, a.part = g.dpart.p, a.out = g.dout.p;
    };
    if (g.prec == MXM_PREC_FP64) {
      SpeedArgs<double> a{};
      a.q = g.ar.qd[0].p;
      fill(a);
      MXM_HIP(launch_max_speed(g.phys, true, &a, g.stream));
    } else {
      SpeedArgs<float> a{};
      a.q = g.ar.qf[0].p;
      fill(a);
      MXM_HIP(launch_max_speed(g.phys, false, &a, g.stream));
    }
    double h = 0.0;
    MXM_HIP(hipMemcpyAsync(&h, g.dout.p, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    MXM_HIP(hipStreamSynchronize(g.stream));
    return h;
  }

  // diagnostics of one instance (Solver::diagnostics over its tiles)
  void diagnostics(int id, double *h) {
    BInst &in = inst(id);
    BGroup &g = settled(id);
    const int nt = in.ntiles(), NC = g.NC;
    const size_t per = (size_t)kDiagSums * 2 + kDiagMaxNC;
    if (g.dpart.n < (size_t)nt * per) g.dpart.alloc((size_t)nt * per);
    if (g.dout.n < (size_t)(3 * NC + 2)) g.dout.alloc(3 * NC + 2);
    auto fill = [&](auto &a) {
      a.invJ = g.ar.invJ.p, a.inv_vol = g.ar.inv_vol.p, a.w = g.w.p, a.Dn = g.Dn.p, a.tile_count = g.ar.tile_count.p;
      a.epb = g.epb, a.tile0 = in.base_tile, a.ntiles = nt, a.NC = NC, a.nvel = diag_nvel(g.phys);
      a.part = g.dpart.p, a.pmax = g.dpart.p + (size_t)nt * kDiagSums * 2, a.out = g.dout.p;
    };
    if (g.prec == MXM_PREC_FP64) {
      DiagArgs<double> a{};
      a.q = g.ar.qd[0].p;
      fill(a);
      MXM_HIP(launch_diagnostics(g.dim, g.P, true, &a, g.stream));
    } else {
      DiagArgs<float> a{};
      a.q = g.ar.qf[0].p;
      fill(a);
      MXM_HIP(launch_diagnostics(g.dim, g.P, false, &a, g.stream));
    }
    MXM_HIP(hipMemcpyAsync(h, g.dout.p, (3 * NC + 2) * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    MXM_HIP(hipStreamSynchronize(g.stream));
    if (!diag_nvel(g.phys)) h[3 * NC] = h[3 * NC + 1] = std::nan("");
  }

  // drops an instance: a waiting one is forgotten; a placed one is
  // switched off on the device (in stream order) and its tiles become a
  // free range; its slot is reused only after the next compaction
  void remove(int id) {
    BInst &in = inst(id);
    BGroup &g = group_with(id);
    auto w = std::find(g.waiting.begin(), g.waiting.end(), id);
    if (w != g.waiting.end()) {
      g.waiting.erase(w);
    } else if (in.placed) {
      hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, g.stream, g.ar.iactive.p + in.slot, 0, 1);
      const int nt = in.ntiles();
      hipLaunchKernelGGL(fill_int_kernel, dim3((nt + 255) / 256), dim3(256), 0, g.stream,
                         g.ar.tile_inst.p + in.base_tile, 0, nt);
      MXM_HIP(hipGetLastError());
      g.fs.give(in.base_tile, nt);
      g.placed.erase(std::find(g.placed.begin(), g.placed.end(), id));
      if (g.placed.empty() && g.waiting.empty()) release(g);
    }
    insts.erase(id);
  }

  // a group with no instances gives its arena back at once
  void release(BGroup &g) {
    drain(g);
    for (auto &e : g.sg) e.reset();
    Arena empty;
    g.ar.swap(empty);
    g.fs = FreeSpace{};
    g.cap_slots = 0, g.next_slot = 1, g.maxj = 1, g.grid_tiles = 0;
    g.calls = g.seen = 0;
  }

  size_t bytes() const {
    size_t b = 0;
    for (auto &kv : groups) b += kv.second->bytes();
    return b;
  }
};

Batch *as_batch(void *h) {
  if (!h) throw ArgError("null batch handle");
  return static_cast<Batch *>(h);
}

int footprint(int dim, const void *desc, int P, int phys, int prec, int limiter, long long *out) {
  if (!desc || !out) throw ArgError("null argument");
  if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
  const int nc = phys_nc(phys);
  if (nc < 0) throw ArgError("unknown physics id");
  RefElem re = build_refelem(dim, P);
  HostMesh m;
  if (dim == 3) {
    auto d = static_cast<const mxm_mesh3d_desc *>(desc);
    check_bc(d->bc, 6);
    m = build_mesh_3d(d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->bc, P, re);
  } else {
    auto d = static_cast<const mxm_mesh2d_desc *>(desc);
    check_bc(d->bc, 4);
    m = build_mesh_2d(d->nx, d->ny, d->lx, d->ly, d->bc, P, re);
  }
  ElemTables t = build_elem_tables(m, re, nc, false);
  *out = (long long)instance_bytes(t, m.NP, nc, m.NFP, dim, prec, limiter != 0);
  return MXM_OK;
}

}  // namespace
}  // namespace mxm

using namespace mxm;

extern "C" {

int mxm_batch_create(int precision, void **out) {
  return guarded([&] {
    if (!out) throw ArgError("null out handle");
    *out = nullptr;
    if (g_device < 0) return fail(MXM_ERR_NOGPU, "mxm_hip_init has not been called");
    if (precision != MXM_PREC_FP64 && precision != MXM_PREC_FP32_MFMA) throw ArgError("unknown precision");
    auto b = std::make_unique<Batch>();
    b->prec = precision;
    MXM_HIP(hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking));
    *out = b.release();
    return MXM_OK;
  });
}

int mxm_batch_autotune(void *h, int on) {
  return guarded([&] {
    as_batch(h)->tune = on != 0;
    return MXM_OK;
  });
}

int mxm_batch_headroom(void *h, double fraction) {
  return guarded([&] {
    if (!(fraction >= 0.0) || fraction > 16.0) throw ArgError("headroom must be in [0, 16]");
    as_batch(h)->headroom = fraction;
    return MXM_OK;
  });
}

int mxm_batch_graphs(void *h, int on) {
  return guarded([&] {
    Batch *b = as_batch(h);
    b->sync_all();
    for (auto &kv : b->groups) {
      kv.second->retire_graphs();
      kv.second->grid_tiles = 0;  // recaptured (or not) at the next step
    }
    b->graphs = on != 0;
    return MXM_OK;
  });
}

int mxm_batch_add_2d(void *h, const mxm_mesh2d_desc *desc, int P, int phys, const double *prm, double dt, int nsteps,
                     int limiter, double eps, const double *q0, long long n, int *id) {
  return guarded([&] {
    if (!desc || !id) throw ArgError("null argument");
    *id = as_batch(h)->add(2, desc, P, phys, prm, dt, nsteps, limiter, eps, q0, n);
    return MXM_OK;
  });
}

int mxm_batch_add_3d(void *h, const mxm_mesh3d_desc *desc, int P, int phys, const double *prm, double dt, int nsteps,
                     int limiter, double eps, const double *q0, long long n, int *id) {
  return guarded([&] {
    if (!desc || !id) throw ArgError("null argument");
    *id = as_batch(h)->add(3, desc, P, phys, prm, dt, nsteps, limiter, eps, q0, n);
    return MXM_OK;
  });
}

int mxm_batch_footprint_2d(const mxm_mesh2d_desc *desc, int P, int phys, int precision, int limiter,
                           long long *bytes) {
  return guarded([&] { return footprint(2, desc, P, phys, precision, limiter, bytes); });
}

int mxm_batch_footprint_3d(const mxm_mesh3d_desc *desc, int P, int phys, int precision, int limiter,
                           long long *bytes) {
  return guarded([&] { return footprint(3, desc, P, phys, precision, limiter, bytes); });
}

int mxm_batch_step(void *h, int nsteps) {
  return guarded([&] {
    if (nsteps < 0) throw ArgError("negative step count");
    as_batch(h)->step(nsteps);
    return MXM_OK;
  });
}

static int poll_into(Batch *b, int lag, int *ids, int cap, int *count) {
  if (!count || (cap > 0 && !ids)) throw ArgError("null argument");
  std::vector<int> done = b->poll(lag);
  if ((int)done.size() > cap) {
    // report what fits; the rest stays unreported for the next poll
    for (size_t k = cap; k < done.size(); ++k) b->inst(done[k]).reported = false;
    done.resize(cap);
  }
  for (size_t k = 0; k < done.size(); ++k) ids[k] = done[k];
  *count = (int)done.size();
  return MXM_OK;
}

int mxm_batch_poll(void *h, int *ids, int cap, int *count) {
  return guarded([&] { return poll_into(as_batch(h), 0, ids, cap, count); });
}

int mxm_batch_poll_lagged(void *h, int *ids, int cap, int *count) {
  return guarded([&] { return poll_into(as_batch(h), 1, ids, cap, count); });
}

int mxm_batch_status(void *h, int id, int *steps_done) {
  return guarded([&] {
    Batch *b = as_batch(h);
    if (!steps_done) throw ArgError("null argument");
    b->sync_all();
    *steps_done = b->inst(id).steps_done;
    return MXM_OK;
  });
}

int mxm_batch_failed(void *h, int id, int *failed) {
  return guarded([&] {
    Batch *b = as_batch(h);
    if (!failed) throw ArgError("null argument");
    *failed = b->inst(id).failed ? 1 : 0;
    return MXM_OK;
  });
}

int mxm_batch_get_state(void *h, int id, double *q, long long n) {
  return guarded([&] {
    as_batch(h)->get_state(id, q, n);
    return MXM_OK;
  });
}

int mxm_batch_max_wave_speed(void *h, int id, double *speed) {
  return guarded([&] {
    if (!speed) throw ArgError("null output");
    *speed = as_batch(h)->max_wave_speed(id);
    return MXM_OK;
  });
}

int mxm_batch_diagnostics(void *h, int id, double *out, int cap, int *n) {
  return guarded([&] {
    Batch *b = as_batch(h);
    const int nv = 3 * b->inst(id).NC + 2;
    if (n) *n = nv;
    if (!out || cap < nv) throw ArgError("diagnostics need 3*NC + 2 values");
    b->diagnostics(id, out);
    return MXM_OK;
  });
}

int mxm_batch_remove(void *h, int id) {
  return guarded([&] {
    as_batch(h)->remove(id);
    return MXM_OK;
  });
}

int mxm_batch_memory(void *h, long long *bytes) {
  return guarded([&] {
    if (!bytes) throw ArgError("null argument");
    *bytes = (long long)as_batch(h)->bytes();
    return MXM_OK;
  });
}

int mxm_batch_sync(void *h) {
  return guarded([&] {
    as_batch(h)->sync_all();
    return MXM_OK;
  });
}

int mxm_batch_destroy(void *h) {
  delete static_cast<Batch *>(h);
  return MXM_OK;
}

}  // extern "C"
