// batch.hip -- the batch engine: many independent DG problems advanced
// together, one stage launch per kernel group and RK stage.
//
// An instance is one problem (mesh, order, physics parameters, time step,
// step count, initial state).  Instances that use the same stage kernel
// (dimension, P, physics, flux variant, limiter) form a group whose tiles
// are concatenated in one device arena: the stage kernel's block reads its
// tile's instance and from it the physics parameters and time step
// (StageArgs::tile_inst / inst_*), so one launch advances every instance
// of the group by one stage.  A device step counter per instance turns the
// instance off (inst_active = 0) after its last step, so instances with
// different step counts share the launches and finished ones cost a block
// exit.  One SSPRK3 step of a group (three stages, limiter, counter bump)
// is captured in a HIP graph and replayed.
//
// Adding or removing instances marks the group dirty; the next step
// repacks it: a new arena in the new member order, the placed states
// copied device to device, new states uploaded, the tables rebuilt.
// Results do not depend on the mix: a tile never mixes instances and every
// element's arithmetic only reads its own instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/mxm_dg.h"
#include "cfl.hpp"
#include "diag.hpp"
#include "limiter.hpp"
#include "runtime.hpp"

namespace mxm {
namespace {

__global__ void bump_steps_kernel(int *steps, const int *nsteps, int *active, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && active[i]) {
    const int s = steps[i] + 1;
    steps[i] = s;
    if (s >= nsteps[i]) active[i] = 0;
  }
}

struct BInst {
  int id = 0, dim = 0, P = 0, phys = 0, variant = 0, NC = 0, K = 0, NP = 0;
  bool limiter = false;
  double eps = 0.1;
  double prm[MXM_PRM_MAX] = {0};
  double dt = 0.0;
  int nsteps = 0, steps_done = 0;
  bool placed = false, reported = false;
  int base_tile = 0;
  std::vector<double> q0;  // initial state, reference order, until placed
  ElemTables t;
};

using GroupKey = std::tuple<int, int, int, int, int, long long>;  // dim, P, phys, variant, limiter, eps bits

// The device arena of a kernel group: state trio, concatenated per-slot and
// per-job tables, per-instance scalars.  A repack builds a whole new arena
// before it replaces the old one, so a failed allocation or copy leaves the
// group as it was.
struct Arena {
  DevBuf<double> qd[3];
  DevBuf<float> qf[3];
  DevBuf<double> mean, invJ, inv_vol, jgeo, iprm, idt, ipost;
  DevBuf<int4> conn, jobs;
  DevBuf<unsigned short> ipr;
  DevBuf<int> tile_count, tile_inst, iactive, isteps, insteps;
  void swap(Arena &o) noexcept {
    for (int k = 0; k < 3; ++k) qd[k].swap(o.qd[k]), qf[k].swap(o.qf[k]);
    mean.swap(o.mean), invJ.swap(o.invJ), inv_vol.swap(o.inv_vol), jgeo.swap(o.jgeo), iprm.swap(o.iprm);
    idt.swap(o.idt), ipost.swap(o.ipost), conn.swap(o.conn), jobs.swap(o.jobs), ipr.swap(o.ipr);
    tile_count.swap(o.tile_count), tile_inst.swap(o.tile_inst), iactive.swap(o.iactive), isteps.swap(o.isteps);
    insteps.swap(o.insteps);
  }
  size_t bytes() const {
    size_t b = 0;
    for (int k = 0; k < 3; ++k) b += qd[k].bytes() + qf[k].bytes();
    return b + mean.bytes() + invJ.bytes() + inv_vol.bytes() + jgeo.bytes() + iprm.bytes() + idt.bytes() +
           ipost.bytes() + conn.bytes() + jobs.bytes() + ipr.bytes() + tile_count.bytes() + tile_inst.bytes() +
           iactive.bytes() + isteps.bytes() + insteps.bytes();
  }
};

struct BGroup {
  int dim = 0, P = 0, phys = 0, variant = 0, NC = 0, NP = 0, NFP = 0, NF = 0, epb = 0, prec = 0;
  int cfg = 0;  // launch configuration (StageCfg), autotuned per packed size
  bool limiter = false;
  double eps = 0.1;
  StageLaunch launch = nullptr;
  std::vector<int> members;  // instance ids in arena order
  std::vector<int> packed;   // the member list the device arena holds
  bool dirty = true;
  int ntiles = 0, maxj = 0;
  // The arena orders the members by their remaining steps at the repack,
  // longest first, so the instances still running always hold a prefix of
  // the tiles.  rem[k] / end_tile[k]: member k's remaining steps and the
  // end of its tiles; launched: steps enqueued since the repack; the graphs
  // launch grid_tiles blocks, shrunk as the running prefix shrinks.
  std::vector<int> rem, end_tile;
  long long launched = 0;
  int grid_tiles = 0;
  // tiles of the members still running after s more steps
  int active_tiles(long long s) const {
    int t = 0;
    for (size_t k = 0; k < rem.size() && rem[k] > s; ++k) t = end_tile[k];
    return t;
  }
  Arena ar;
  DevBuf<double> Dt, Lt, w;
  DevBuf<double> Dn, dpart, dout;  // diagnostics
  DevBuf<float> Aop;
  DevBuf<int> fmap;
  // replayed HIP graphs of kGraphSteps steps ([0]) and of one step ([1])
  static constexpr int kGraphSteps = 8;
  StepGraph sg[2];
  GroupKey key{};

  ~BGroup() { drop_graph(); }
  void drop_graph() {
    for (auto &e : sg) e.reset();
  }
  size_t bytes() const {
    return ar.bytes() + Dn.bytes() + dpart.bytes() + dout.bytes() + Dt.bytes() + Lt.bytes() + w.bytes() +
           Aop.bytes() + fmap.bytes();
  }
};

// Device bytes one instance adds to its group's arena (exact: the same
// tables the repack builds).
size_t instance_bytes(const ElemTables &t, int NP, int NC, int NFP, int dim, int prec, bool limiter) {
  const size_t slots = (size_t)t.til.slots(), tiles = (size_t)t.til.ntiles, js = tiles * t.fj.maxj;
  const size_t word = prec == MXM_PREC_FP64 ? sizeof(double) : sizeof(float);
  size_t b = 3 * slots * NP * NC * word;                            // state trio
  b += slots * (dim * dim + 1) * sizeof(double);                    // invJ, inv_vol
  b += slots * (dim + 1) * sizeof(int4);                            // limiter neighbours
  b += js * (sizeof(int4) + 4 * sizeof(double) + NFP * sizeof(unsigned short));  // jobs, geometry, partner nodes
  b += tiles * 2 * sizeof(int);                                     // tile_count, tile_inst
  b += (MXM_PRM_MAX + 2) * sizeof(double) + 3 * sizeof(int);        // per-instance scalars
  if (limiter) b += slots * NC * sizeof(double);                    // cell means
  return b;
}

struct Batch {
  int prec = MXM_PREC_FP64;
  hipStream_t stream = nullptr;
  std::map<GroupKey, std::unique_ptr<BGroup>> groups;
  std::map<int, std::unique_ptr<BInst>> insts;
  std::map<std::pair<int, int>, RefElem> refs;
  int next_id = 0;
  // autotuned launch configuration per (group, tile-count octave)
  bool tune = true;
  std::map<std::pair<GroupKey, int>, int> tuned;

  ~Batch() {
    groups.clear();
    if (stream) (void)hipStreamDestroy(stream);
  }

  const RefElem &ref(int dim, int P) {
    auto k = std::make_pair(dim, P);
    auto it = refs.find(k);
    if (it == refs.end()) it = refs.emplace(k, build_refelem(dim, P)).first;
    return it->second;
  }

  BGroup &group_of(const BInst &in) {
    long long eb = 0;
    std::memcpy(&eb, &in.eps, sizeof(eb));
    GroupKey key{in.dim, in.P, in.phys, in.variant, in.limiter ? 1 : 0, in.limiter ? eb : 0};
    auto it = groups.find(key);
    if (it != groups.end()) return *it->second;
    auto g = std::make_unique<BGroup>();
    const RefElem &re = ref(in.dim, in.P);
    g->dim = in.dim, g->P = in.P, g->phys = in.phys, g->variant = in.variant, g->NC = in.NC, g->NP = re.NP,
    g->NFP = re.NFP, g->NF = re.NF;
    g->epb = stage_epb(in.dim, re.NP, re.NFP, in.NC);
    g->prec = prec;
    g->limiter = in.limiter, g->eps = in.eps;
    g->key = key;
    g->cfg = stage_cfg_default(prec);
    g->launch = find_stage(in.dim, in.phys, in.variant, in.P, prec, g->cfg);
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
    return *groups.emplace(key, std::move(g)).first->second;
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
    in->t = build_elem_tables(m, re, nc);
    in->id = next_id++;
    BGroup &g = group_of(*in);
    g.members.push_back(in->id);
    g.dirty = true;
    const int id = in->id;
    insts.emplace(id, std::move(in));
    return id;
  }

  BInst &inst(int id) {
    auto it = insts.find(id);
    if (it == insts.end()) throw ArgError("unknown instance id");
    return *it->second;
  }

  BGroup &group_with(int id) {
    for (auto &kv : groups)
      for (int m : kv.second->members)
        if (m == id) return *kv.second;
    throw ArgError("instance has no group");
  }

  // host mirrors of the device step counters of the packed arena
  void pull_counters(BGroup &g) {
    if (g.packed.empty() || !g.ar.isteps.p) return;
    std::vector<int> st(g.packed.size());
    MXM_HIP(hipMemcpy(st.data(), g.ar.isteps.p, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < g.packed.size(); ++k) {
      auto it = insts.find(g.packed[k]);
      if (it != insts.end() && it->second->placed) it->second->steps_done = st[k];
    }
  }

  // The new arena's state trio: placed members' tiles copied from the old
  // arena (their old base tiles are still in BInst::base_tile), new members
  // uploaded from q0.
  template <class R>
  void place_states(const BGroup &g, DevBuf<R> (&q)[3], DevBuf<R> (&nq)[3], const std::vector<int> &order,
                    const std::vector<int> &new_base) {
    const size_t tile_vals = (size_t)g.epb * g.NP * g.NC;
    for (auto &b : nq) {
      b.alloc(std::max<size_t>(1, (size_t)g.ntiles * tile_vals));
      MXM_HIP(hipMemset(b.p, 0, b.bytes()));
    }
    for (size_t k = 0; k < order.size(); ++k) {
      const BInst &in = inst(order[k]);
      R *dst = nq[0].p + (size_t)new_base[k] * tile_vals;
      const size_t nv = (size_t)in.t.til.ntiles * tile_vals;
      if (in.placed) {
        MXM_HIP(hipMemcpy(dst, q[0].p + (size_t)in.base_tile * tile_vals, nv * sizeof(R), hipMemcpyDeviceToDevice));
      } else {
        const size_t node = (size_t)g.NP * g.NC;
        std::vector<R> d(nv, R(0));
        for (int e = 0; e < in.K; ++e)
          for (size_t c = 0; c < node; ++c) d[(size_t)in.t.til.dev_of[e] * node + c] = R(in.q0[(size_t)e * node + c]);
        MXM_HIP(hipMemcpy(dst, d.data(), nv * sizeof(R), hipMemcpyHostToDevice));
      }
    }
  }

  // Rebuilds the group's arena for its current member list.  Everything is
  // built in a new arena first; only when every allocation and copy has
  // succeeded are the arena, the member order, the base tiles and the
  // placed flags replaced, so an error leaves the group as it was (still
  // dirty, to be retried).
  void repack(BGroup &g) {
    MXM_HIP(hipStreamSynchronize(stream));
    pull_counters(g);
    g.drop_graph();
    std::vector<int> order = g.members;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const BInst &x = inst(a), &y = inst(b);
      return x.nsteps - x.steps_done > y.nsteps - y.steps_done;
    });
    std::vector<int> new_base(order.size());
    int tiles = 0, maxj = 1;
    for (size_t k = 0; k < order.size(); ++k) {
      const BInst &in = inst(order[k]);
      new_base[k] = tiles;
      tiles += in.t.til.ntiles;
      maxj = std::max(maxj, in.t.fj.maxj);
    }
    const int old_ntiles = g.ntiles;
    g.ntiles = tiles;  // place_states sizes the new trio from it
    Arena na;
    try {
      if (g.prec == MXM_PREC_FP64)
        place_states(g, g.ar.qd, na.qd, order, new_base);
      else
        place_states(g, g.ar.qf, na.qf, order, new_base);
    } catch (...) {
      g.ntiles = old_ntiles;
      throw;
    }
    g.ntiles = old_ntiles;
    // concatenated tables, slots offset by each instance's base
    const int epb = g.epb, dd = g.dim * g.dim, NF = g.NF, NFP = g.NFP;
    const size_t slots = (size_t)tiles * epb;
    std::vector<double> ij(slots * dd, 0.0), iv(slots, 0.0), geo((size_t)tiles * maxj * 4, 0.0);
    std::vector<int4> cn(slots * NF), jb((size_t)tiles * maxj);
    std::vector<unsigned short> pr((size_t)tiles * maxj * NFP, 0);
    std::vector<int> tc(tiles, 0), ti(tiles, 0);
    const size_t ni = order.size();
    std::vector<double> iprm(ni * MXM_PRM_MAX), idt(ni), ipost(ni);
    std::vector<int> iact(ni), ist(ni), inst_n(ni);
    for (size_t k = 0; k < ni; ++k) {
      const BInst &in = inst(order[k]);
      const ElemTables &t = in.t;
      const int bt = new_base[k], bs = bt * epb, lt = t.til.ntiles, lmj = t.fj.maxj;
      for (int s = 0; s < lt * epb; ++s) {
        for (int q = 0; q < dd; ++q) ij[(size_t)(bs + s) * dd + q] = t.invJ[(size_t)s * dd + q];
        iv[bs + s] = t.inv_vol[s];
        for (int lf = 0; lf < NF; ++lf) {
          int4 c = t.conn[(size_t)s * NF + lf];
          c.x += bs;
          cn[(size_t)(bs + s) * NF + lf] = c;
        }
      }
      for (int T = 0; T < lt; ++T) {
        tc[bt + T] = t.til.count[T];
        ti[bt + T] = (int)k;
        for (int j = 0; j < maxj; ++j) {
          const size_t dj = (size_t)(bt + T) * maxj + j;
          if (j < lmj) {
            const size_t sj = (size_t)T * lmj + j;
            int4 r{t.fj.jobs[sj * 4], t.fj.jobs[sj * 4 + 1] + bs, t.fj.jobs[sj * 4 + 2], t.fj.jobs[sj * 4 + 3]};
            jb[dj] = r;
            for (int c = 0; c < 4; ++c) geo[dj * 4 + c] = t.fj.geo[sj * 4 + c];
            for (int r2 = 0; r2 < NFP; ++r2) pr[dj * NFP + r2] = t.fj.ipr[sj * NFP + r2];
          } else {
            jb[dj] = int4{-1, (bt + T) * epb, 0, 0};
          }
        }
      }
      std::memcpy(&iprm[k * MXM_PRM_MAX], in.prm, sizeof(in.prm));
      idt[k] = in.dt;
      ipost[k] = post_factor(in.phys, in.prm, in.dt);
      ist[k] = in.steps_done;
      inst_n[k] = in.nsteps;
      iact[k] = in.steps_done < in.nsteps ? 1 : 0;
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
    if (g.limiter) na.mean.alloc(slots * g.NC);
    // ---- commit (no more allocations or copies that can fail) ----
    g.ar.swap(na);
    g.ntiles = tiles;
    g.maxj = maxj;
    g.members = order;
    g.packed = order;
    for (size_t k = 0; k < ni; ++k) {
      BInst &in = inst(order[k]);
      in.base_tile = new_base[k];
      if (!in.placed) {
        in.placed = true;
        in.q0.clear();
        in.q0.shrink_to_fit();
      }
    }
    g.dirty = false;
    g.rem.resize(ni);
    g.end_tile.resize(ni);
    for (size_t k = 0; k < ni; ++k) {
      const BInst &in = inst(order[k]);
      g.rem[k] = std::max(0, in.nsteps - in.steps_done);
      g.end_tile[k] = in.base_tile + in.t.til.ntiles;
    }
    g.launched = 0;
    recapture(g, g.active_tiles(0));
  }

  // graphs over the first `tiles` tiles of the arena (none when 0); the
  // old graphs finish before they are destroyed
  void recapture(BGroup &g, int tiles) {
    MXM_HIP(hipStreamSynchronize(stream));
    g.drop_graph();
    g.grid_tiles = tiles;
    if (tiles > 0) {
      if (tune) choose_config(g);
      capture(g);
    }
  }

  // the launch configuration for the group's packed size: from the cache
  // of this (group, tile-count octave), else timed on stage 1 (q -> q1,
  // which the next step rewrites before reading)
  void choose_config(BGroup &g) {
    int oct = 0;
    while ((1 << oct) < g.grid_tiles) ++oct;
    const auto key = std::make_pair(g.key, oct);
    auto it = tuned.find(key);
    if (it == tuned.end()) {
      MXM_HIP(hipStreamSynchronize(stream));
      int c;
      if (g.prec == MXM_PREC_FP64) {
        const StageArgs<double> a = stage_args(g, g.ar.qd, 0);
        c = autotune_stage(g.dim, g.phys, g.variant, g.P, g.prec, &a, stream, 10, nullptr);
      } else {
        const StageArgs<float> a = stage_args(g, g.ar.qf, 0);
        c = autotune_stage(g.dim, g.phys, g.variant, g.P, g.prec, &a, stream, 10, nullptr);
      }
      MXM_HIP(hipStreamSynchronize(stream));
      it = tuned.emplace(key, c).first;
    }
    g.cfg = it->second;
    g.launch = find_stage(g.dim, g.phys, g.variant, g.P, g.prec, g.cfg);
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
  void launch_step(BGroup &g, DevBuf<R> (&q)[3]) {
    const int ni = (int)g.members.size();
    for (int s = 0; s < 3; ++s) {
      const StageArgs<R> a = stage_args(g, q, s);
      MXM_HIP(g.launch(&a, stream));
      if (g.limiter)
        MXM_HIP(launch_limiter<R>(a.q_out, g.ar.mean.p, g.w.p, g.ar.conn.p, g.grid_tiles * g.epb, g.NP, g.NC, g.NF,
                                  g.dim, g.eps, stream, g.ar.tile_inst.p, g.ar.iactive.p, g.epb));
    }
    hipLaunchKernelGGL(bump_steps_kernel, dim3((ni + 255) / 256), dim3(256), 0, stream, g.ar.isteps.p,
                       g.ar.insteps.p, g.ar.iactive.p, ni);
    MXM_HIP(hipGetLastError());
  }

  void launch_step(BGroup &g) {
    if (g.prec == MXM_PREC_FP64)
      launch_step(g, g.ar.qd);
    else
      launch_step(g, g.ar.qf);
  }

  // n SSPRK3 steps of the group as one HIP graph (kernel nodes only); the
  // graph is kept beside its executable
  void capture_steps(BGroup &g, int n, StepGraph &out) {
    MXM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < n; ++k) launch_step(g);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    MXM_HIP(hipStreamEndCapture(stream, &out.graph));
    MXM_HIP(hipGraphInstantiate(&out.exec, out.graph, nullptr, nullptr, 0));
  }
  void capture(BGroup &g) {
    capture_steps(g, BGroup::kGraphSteps, g.sg[0]);
    capture_steps(g, 1, g.sg[1]);
  }

  void step(int n) {
    for (auto &kv : groups)
      if (kv.second->dirty) repack(*kv.second);
    // instances that finish inside a launch are switched off by their
    // device counters; the grid shrinks to the running prefix between
    // launches once it has lost an eighth of its tiles
    auto launch_all = [&](int which, int k) {
      for (auto &kv : groups) {
        BGroup &g = *kv.second;
        const int need = g.active_tiles(g.launched);
        g.launched += k;
        if (need == 0) continue;
        if (need <= g.grid_tiles - g.grid_tiles / 8) recapture(g, need);
        MXM_HIP(hipGraphLaunch(g.sg[which].exec, stream));
      }
    };
    for (; n >= BGroup::kGraphSteps; n -= BGroup::kGraphSteps) launch_all(0, BGroup::kGraphSteps);
    for (; n > 0; --n) launch_all(1, 1);
  }

  void sync_counters() {
    MXM_HIP(hipStreamSynchronize(stream));
    for (auto &kv : groups) pull_counters(*kv.second);
  }

  template <class R>
  void read_state(BGroup &g, DevBuf<R> (&q)[3], BInst &in, double *h) {
    const size_t node = (size_t)g.NP * g.NC, tile_vals = (size_t)g.epb * node;
    const size_t nv = (size_t)in.t.til.ntiles * tile_vals;
    std::vector<R> d(nv);
    MXM_HIP(hipMemcpy(d.data(), q[0].p + (size_t)in.base_tile * tile_vals, nv * sizeof(R), hipMemcpyDeviceToHost));
    for (int e = 0; e < in.K; ++e)
      for (size_t c = 0; c < node; ++c) h[(size_t)e * node + c] = double(d[(size_t)in.t.til.dev_of[e] * node + c]);
  }

  void get_state(int id, double *h, long long n) {
    BInst &in = inst(id);
    if (!h || n != (long long)in.K * in.NP * in.NC) throw ArgError("state length must be K*NP*NC");
    if (!in.placed) {
      std::memcpy(h, in.q0.data(), n * sizeof(double));
      return;
    }
    BGroup &g = group_with(id);
    if (g.dirty) repack(g);
    MXM_HIP(hipStreamSynchronize(stream));
    if (g.prec == MXM_PREC_FP64)
      read_state(g, g.ar.qd, in, h);
    else
      read_state(g, g.ar.qf, in, h);
  }

  // the largest signal speed over the nodes of one instance
  double max_wave_speed(int id) {
    BInst &in = inst(id);
    BGroup &g = group_with(id);
    if (g.dirty) repack(g);
    const int nt = in.t.til.ntiles;
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
      MXM_HIP(launch_max_speed(g.phys, true, &a, stream));
    } else {
      SpeedArgs<float> a{};
      a.q = g.ar.qf[0].p;
      fill(a);
      MXM_HIP(launch_max_speed(g.phys, false, &a, stream));
    }
    double h = 0.0;
    MXM_HIP(hipMemcpyAsync(&h, g.dout.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    MXM_HIP(hipStreamSynchronize(stream));
    return h;
  }

  // diagnostics of one instance (Solver::diagnostics over its tiles)
  void diagnostics(int id, double *h) {
    BInst &in = inst(id);
    BGroup &g = group_with(id);
    if (g.dirty) repack(g);
    const int nt = in.t.til.ntiles, NC = g.NC;
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
      MXM_HIP(launch_diagnostics(g.dim, g.P, true, &a, stream));
    } else {
      DiagArgs<float> a{};
      a.q = g.ar.qf[0].p;
      fill(a);
      MXM_HIP(launch_diagnostics(g.dim, g.P, false, &a, stream));
    }
    MXM_HIP(hipMemcpyAsync(h, g.dout.p, (3 * NC + 2) * sizeof(double), hipMemcpyDeviceToHost, stream));
    MXM_HIP(hipStreamSynchronize(stream));
    if (!diag_nvel(g.phys)) h[3 * NC] = h[3 * NC + 1] = std::nan("");
  }

  void remove(int id) {
    BInst &in = inst(id);
    (void)in;
    BGroup &g = group_with(id);
    MXM_HIP(hipStreamSynchronize(stream));
    pull_counters(g);
    g.members.erase(std::find(g.members.begin(), g.members.end(), id));
    g.dirty = true;
    insts.erase(id);
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
  ElemTables t = build_elem_tables(m, re, nc);
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
    MXM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
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

int mxm_batch_poll(void *h, int *ids, int cap, int *count) {
  return guarded([&] {
    Batch *b = as_batch(h);
    if (!count || (cap > 0 && !ids)) throw ArgError("null argument");
    b->sync_counters();
    int k = 0;
    for (auto &kv : b->insts) {
      BInst &in = *kv.second;
      if (!in.reported && in.steps_done >= in.nsteps && k < cap) {
        ids[k++] = in.id;
        in.reported = true;
      }
    }
    *count = k;
    return MXM_OK;
  });
}

int mxm_batch_status(void *h, int id, int *steps_done) {
  return guarded([&] {
    Batch *b = as_batch(h);
    if (!steps_done) throw ArgError("null argument");
    b->sync_counters();
    *steps_done = b->inst(id).steps_done;
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
    MXM_HIP(hipStreamSynchronize(as_batch(h)->stream));
    return MXM_OK;
  });
}

int mxm_batch_destroy(void *h) {
  delete static_cast<Batch *>(h);
  return MXM_OK;
}

}  // extern "C"
