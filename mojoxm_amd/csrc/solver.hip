// solver.hip -- the C ABI of libmxm_hip.so (include/mxm_dg.h): host-side
// builders, the device-resident solver (Solver[PhysT, P] of
// src/solver.mojo:538-1093 and the 2D <phys>_rk_stage_2d drivers) and its
// SSPRK3 stepping, limiter, throughput and memory reports.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mxm_dg.h"
#include "cfl.hpp"
#include "diag.hpp"
#include "host.hpp"
#include "limiter.hpp"
#include "runtime.hpp"

namespace mxm {
namespace {

// Scratch of mxm_dg_max_wave_speed_noalloc: the per-tile maxima of the
// speed reduction, one double per tile, for at most kNoAllocTiles tiles
// (2 MiB).  Every device holds its own copy for the life of the process;
// its address on a device comes from hipGetSymbolAddress with that device
// current (mxm_hip_init), and host code uses only that address
// (g_scratch), never the array's name.  All calls on one device share it:
// two calls that overlap on one device race.
constexpr int kNoAllocTiles = 1 << 18;
static __device__ __align__(256) double scratch[kNoAllocTiles];
constexpr int kMaxDevices = 64;
// MI355X published peaks: HBM3E bandwidth (GB/s) and the dense FP32
// matrix rate of V_MFMA_F32_* (TFLOP/s)
constexpr double kHbmPeakGBs = 8000.0, kMfmaF32PeakTFs = 157.3;
double *g_scratch[kMaxDevices] = {};

struct Solver {
  int device = 0;  // the device current at create time
  int dim = 0, P = 0, NP = 0, NFP = 0, NF = 0, NC = 0, phys = 0;
  int K = 0, Nf = 0;
  double prm[MXM_PRM_MAX] = {0};
  int prec = MXM_PREC_FP64;
  int variant = 0, cfg = 0;  // flux variant; launch configuration (StageCfg)
  // wide tiles (stage_epb): the FP32-MFMA form of a shape that has them,
  // chosen by autotune() or set by mxm_dg_tile_config; the host mesh is
  // kept to lay the tables out again
  bool wide = false;
  std::unique_ptr<HostMesh> hmesh;
  std::unique_ptr<RefElem> hre;
  StageLaunch launch = nullptr;
  hipStream_t stream = nullptr;
  bool limiter = false;
  double eps = 0.1;
  // state trio q, q1, q2: FP64 (MXM_PREC_FP64) or FP32 (MXM_PREC_FP32_MFMA)
  DevBuf<double> q, q1, q2;
  DevBuf<float> f, f1, f2;
  // limiter: cell means [3][Kd*NC] and factors [3][Kd] of the stage
  // buffers (q, q1, q2), node extrema [Kd*NCL][2] of the last stage output
  // (the limited components: all NC in 3D, component 0 in 2D)
  DevBuf<double> lmean, ltheta, lext;
  DevBuf<double> invJ, inv_vol, jgeo, Dt, Lt, w, dprm;
  DevBuf<double> Dn, dpart, dout;  // diagnostics: nodal derivative, tile partials, results
  DevBuf<float> Aop;
  DevBuf<int> conn;  // limiter: partner slot per face
  DevBuf<int> fmap, tile_count;
  DevBuf<int4> jobs;
  DevBuf<unsigned short> ipr;
  int maxj = 0;
  // Device element order: tiles of lattice cells, each contiguous
  // (slot T*epb + el); the host side of the ABI keeps the reference order.
  Tiling til;
  int Kd = 0;  // device element slots, padding included
  int ntiles = 0;

  // HIP graphs of SSPRK3 steps (kernel nodes only), replayed by step(): one
  // of kGraphSteps steps and one of exactly r steps for each remainder
  // r = n mod kGraphSteps asked for, so step(n) is at most n / kGraphSteps
  // + 1 graph launches.  A graph bakes in dt, the limiter setting and the
  // launch configuration; all are rebuilt when any of them changed.  A
  // replayed graph costs about 1 us less per kernel than stream launches of
  // the same kernels, and one launch of many steps keeps the host out of
  // the way of short runs.
  static constexpr int kGraphSteps = 32;
  StepGraph gfull;
  std::map<int, StepGraph> grem;
  double g_dt = 0.0, g_eps = 0.0;
  bool g_lim = false;
  int g_cfg = -1;

  ~Solver() {
    drop_graphs();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graphs() {
    gfull.reset();
    for (auto &kv : grem) kv.second.reset();
    grem.clear();
  }

  // n SSPRK3 steps as one HIP graph, kept beside its executable
  void capture_steps(double dt, int n, StepGraph &out) {
    MXM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < n; ++k)
        for (int st = 0; st < 3; ++st) stage(st, dt);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    MXM_HIP(hipStreamEndCapture(stream, &out.graph));
    MXM_HIP(hipGraphInstantiate(&out.exec, out.graph, nullptr, nullptr, 0));
    // the executable goes to the device now, not at its first launch
    MXM_HIP(hipGraphUpload(out.exec, stream));
  }

  void build(HostMesh m, const RefElem &re) {
    dim = m.dim, P = m.P, NP = m.NP, NFP = m.NFP, NF = m.NF, K = m.K, Nf = m.Nf;
    MXM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    layout(m, re);
    OpTables o = build_operators(re, dim, NC, prec);
    if (prec == MXM_PREC_FP64) {
      Dt.upload(o.Dt.data(), o.Dt.size());
      Lt.upload(o.Lt.data(), o.Lt.size());
    } else {
      Aop.upload(o.Aop.data(), o.Aop.size());
    }
    w.upload(re.w.data(), re.w.size());
    Dn.upload(re.Dn.data(), re.Dn.size());
    fmap.upload(re.fmap.data(), re.fmap.size());
    dprm.upload(prm, MXM_PRM_MAX);
    MXM_HIP(hipDeviceSynchronize());
    if (has_wide()) {
      hmesh = std::make_unique<HostMesh>(std::move(m));
      hre = std::make_unique<RefElem>(re);
    }
  }

  bool has_wide() const { return prec == MXM_PREC_FP32_MFMA && stage_has_wide(dim, NP, NFP, NC); }
  StageLaunch find(int c, bool lim) const { return find_stage(dim, phys, variant, P, prec, c, lim, wide); }

  // the tables and buffers that depend on the tiling (tile size `wide`),
  // with a zero state and nothing pending; graphs are dropped
  void layout(const HostMesh &m, const RefElem &re) {
    drop_graphs();
    ElemTables t = build_elem_tables(m, re, NC, wide);
    til = std::move(t.til);
    Kd = til.slots();
    ntiles = til.ntiles;
    const size_t nq = (size_t)Kd * NP * NC;
    if (prec == MXM_PREC_FP64) {
      for (DevBuf<double> *b : {&q, &q1, &q2}) {
        b->alloc(nq);
        MXM_HIP(hipMemset(b->p, 0, b->bytes()));
      }
    } else {
      for (DevBuf<float> *b : {&f, &f1, &f2}) {
        b->alloc(nq);
        MXM_HIP(hipMemset(b->p, 0, b->bytes()));
      }
    }
    lmean.alloc((size_t)3 * Kd * NC);
    lext.alloc((size_t)2 * Kd * (dim == 3 ? NC : 1));  // the limited components only
    ltheta.alloc((size_t)3 * Kd);
    MXM_HIP(launch_fill(ltheta.p, (long long)ltheta.n, 1.0, stream));
    invJ.upload(t.invJ.data(), t.invJ.size());
    inv_vol.upload(t.inv_vol.data(), t.inv_vol.size());
    conn.upload(t.conn.data(), t.conn.size());
    maxj = t.fj.maxj;
    tile_count.upload(til.count.data(), til.count.size());
    jobs.upload(reinterpret_cast<const int4 *>(t.fj.jobs.data()), t.fj.jobs.size() / 4);
    jgeo.upload(t.fj.geo.data(), t.fj.geo.size());
    ipr.upload(t.fj.ipr.data(), t.fj.ipr.size());
    MXM_HIP(hipDeviceSynchronize());
  }

  // Untimed steps (stream launches) until about `ms` milliseconds of them
  // have run: after an idle gap (a state upload, host work) the device
  // clocks up over 20-30 ms of work, and steps timed sooner read up to 20 %
  // slower (the 32^3 Taylor-Green stage: 224 us right after an upload, 196
  // after 17 ms of steps, 191 after 60 ms).
  static constexpr double kWarmUpMs = 30.0;
  void warm_up(double dt, double ms) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 1, total = 0; total < 65536; total += k, k = k < 256 ? 2 * k : 256) {
      for (int j = 0; j < k; ++j)
        for (int st = 0; st < 3; ++st) stage(st, dt);
      MXM_HIP(hipStreamSynchronize(stream));
      if (std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() >= ms) break;
    }
    if (limiter) materialize(0);
  }

  // moves the state to tiles of size w (its values unchanged)
  void set_tile(bool w) {
    if (w == wide) return;
    if (w && !has_wide()) throw ArgError("this shape and form have no wide tile");
    if (limiter) materialize(0);
    std::vector<double> h((size_t)dof());
    download_state(h.data());
    const bool w0 = wide;
    try {
      wide = w;
      layout(*hmesh, *hre);
      upload_state(h.data());
    } catch (...) {
      // back to the former tiling with the same state, then report
      wide = w0;
      layout(*hmesh, *hre);
      upload_state(h.data());
      launch = find(cfg, false);
      throw;
    }
    launch = find(cfg, false);
  }

  // Shu-Osher plan of src/ssprk3.mojo:41-78 (FP64 coefficients).
  void stage(int s, double dt) {
    if (prec == MXM_PREC_FP64)
      stage_with(s, dt, q.p, q1.p, q2.p);
    else
      stage_with(s, dt, f.p, f1.p, f2.p);
  }

  template <class R>
  StageArgs<R> stage_args(int s, double dt, R *b0, R *b1, R *b2) const {
    StageArgs<R> a{};
    R *trio[3] = {b0, b1, b2};
    const StagePlan pl = stage_plan(s);
    a.q_in = trio[pl.in], a.q_a = trio[pl.qa], a.q_out = trio[pl.out];
    a.a = pl.a, a.b = pl.b, a.c = pl.c;
    a.inst_prm = dprm.p;  // one instance: tile_inst, inst_dt, inst_post, inst_active stay null
    a.dt = dt;
    a.tile_count = tile_count.p;
    a.invJ = invJ.p, a.inv_vol = inv_vol.p;
    a.Dt = Dt.p, a.Lt = Lt.p, a.Aop = Aop.p, a.fmap = fmap.p;
    a.jobs = jobs.p, a.jgeo = jgeo.p, a.ipr = ipr.p, a.maxj = maxj, a.ntiles = ntiles;
    a.post_comp = s == 2 ? post_component(phys, prm) : -1;
    a.post_scale = post_factor(phys, prm, dt);
    // limiter: the stage epilogue writes the output's cell means and node
    // extrema, the theta kernel its factors; the inputs' pending factors
    // are applied as they are read
    if (limiter) {
      const size_t kn = (size_t)Kd * NC;
      a.mean = lmean.p + pl.out * kn, a.ext = lext.p;
      a.lim_th_in = ltheta.p + (size_t)pl.in * Kd, a.lim_mn_in = lmean.p + pl.in * kn;
      a.lim_th_a = ltheta.p + (size_t)pl.qa * Kd, a.lim_mn_a = lmean.p + pl.qa * kn;
    }
    a.w = w.p;
    return a;
  }

  template <class R>
  void stage_with(int s, double dt, R *b0, R *b1, R *b2) {
    const StageArgs<R> a = stage_args(s, dt, b0, b1, b2);
    MXM_HIP(kernel()(&a, stream));
    if (limiter)
      MXM_HIP(launch_bj_theta(ltheta.p + (size_t)stage_plan(s).out * Kd, a.mean, lext.p, conn.p, Kd, NC, NF, dim,
                              eps, stream, tile_count.p, til.epb));
  }

  // the stage kernel of the launch configuration, in its limited variant
  // when the limiter is on
  StageLaunch kernel() const { return limiter ? find(cfg, true) : launch; }

  // writes the limited values of stage buffer b (0 = q) and clears its
  // pending factors
  void materialize(int b) {
    const size_t kn = (size_t)Kd * NC;
    if (prec == MXM_PREC_FP64) {
      double *t[3] = {q.p, q1.p, q2.p};
      MXM_HIP(launch_bj_apply<double>(t[b], ltheta.p + (size_t)b * Kd, lmean.p + b * kn, Kd, NP, NC, stream));
    } else {
      float *t[3] = {f.p, f1.p, f2.p};
      MXM_HIP(launch_bj_apply<float>(t[b], ltheta.p + (size_t)b * Kd, lmean.p + b * kn, Kd, NP, NC, stream));
    }
  }

  // times every launch configuration of the current tiling on SSPRK3
  // steps of the current state (a warm-up; then per configuration one
  // untimed step and reps steps between two events on the solver's
  // stream, stream launches), keeps the fastest and returns its time per
  // stage.  The state is saved to the host first and restored after, also
  // when a call fails.  Full steps, rather than one stage repeated on fixed
  // inputs, see the cache state of the real cycle: on a state past the
  // Infinity Cache the repeated stage ranks the configurations differently.
  double tune(double dt, int reps, double *us) {
    if (limiter) materialize(0);
    std::vector<double> h((size_t)dof());
    download_state(h.data());
    const int c0 = cfg;
    try {
      const double t = time_configs(dt, reps, us);
      upload_state(h.data());
      return t;
    } catch (...) {
      cfg = c0;
      launch = find(cfg, false);
      upload_state(h.data());
      throw;
    }
  }
  double time_configs(double dt, int reps, double *us) {
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Events {
      hipEvent_t *e;
      ~Events() {
        for (int k = 0; k < 2; ++k)
          if (e[k]) (void)hipEventDestroy(e[k]);
      }
    } own{ev};
    for (auto &e : ev) MXM_HIP(hipEventCreate(&e));
    int best = stage_cfg_default(prec);
    double tbest = 0.0;
    // a device coming out of idle runs the first steps slower, which would
    // favour the configurations timed last
    warm_up(dt, kWarmUpMs);
    for (int c = 0; c < stage_cfg_count(prec); ++c) {
      cfg = c;
      launch = find(c, false);
      if (!launch) throw ArgError("no stage kernel for a launch configuration");
      for (int st = 0; st < 3; ++st) stage(st, dt);
      MXM_HIP(hipEventRecord(ev[0], stream));
      for (int r = 0; r < reps; ++r)
        for (int st = 0; st < 3; ++st) stage(st, dt);
      MXM_HIP(hipEventRecord(ev[1], stream));
      MXM_HIP(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      MXM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      const double t = (double)ms * 1e3 / (3.0 * reps);
      if (us) us[c] = t;
      if (c == 0 || t < tbest) best = c, tbest = t;
    }
    cfg = best;
    launch = find(cfg, false);
    return tbest;
  }

  // the fastest launch configuration, and for a shape with a wide tile
  // the faster tiling: the state moves to the other tiling, its
  // configurations are timed, and it moves back if none is faster
  int autotune(double dt, int reps, double *us) {
    const double t0 = tune(dt, reps, us);
    if (!has_wide() || !hmesh) return cfg;
    const int c0 = cfg;
    double u1[kStageCfgMax] = {0};
    set_tile(!wide);
    const double t1 = tune(dt, reps, u1);
    if (t1 < t0) {
      if (us)
        for (int k = 0; k < stage_cfg_count(prec); ++k) us[k] = u1[k];
      return cfg;
    }
    cfg = c0;
    set_tile(!wide);
    return cfg;
  }

  // int q_c, int q_c^2, max |q_c|, KE, enstrophy of the current state
  // (3*NC + 2 values; NaN for KE and enstrophy without a velocity)
  void diagnostics(double *h) {
    const size_t per = (size_t)kDiagSums * 2 + kDiagMaxNC;
    if (dpart.n < (size_t)ntiles * per) dpart.alloc((size_t)ntiles * per);
    if (dout.n < (size_t)(3 * NC + 2)) dout.alloc(3 * NC + 2);
    auto fill = [&](auto &a) {
      a.invJ = invJ.p, a.inv_vol = inv_vol.p, a.w = w.p, a.Dn = Dn.p, a.tile_count = tile_count.p;
      a.epb = til.epb, a.tile0 = 0, a.ntiles = ntiles, a.NC = NC, a.nvel = diag_nvel(phys);
      a.part = dpart.p, a.pmax = dpart.p + (size_t)ntiles * kDiagSums * 2, a.out = dout.p;
    };
    if (prec == MXM_PREC_FP64) {
      DiagArgs<double> a{};
      a.q = q.p;
      fill(a);
      MXM_HIP(launch_diagnostics(dim, P, true, &a, stream));
    } else {
      DiagArgs<float> a{};
      a.q = f.p;
      fill(a);
      MXM_HIP(launch_diagnostics(dim, P, false, &a, stream));
    }
    MXM_HIP(hipMemcpyAsync(h, dout.p, (3 * NC + 2) * sizeof(double), hipMemcpyDeviceToHost, stream));
    MXM_HIP(hipStreamSynchronize(stream));
    if (!diag_nvel(phys)) h[3 * NC] = h[3 * NC + 1] = std::nan("");
  }

  // enqueues the largest-signal-speed reduction of the current state on
  // the solver's stream: tile maxima into part[ntiles], the maximum into
  // out[0] (both device memory)
  void enqueue_max_speed(double *part, double *out) {
    auto fill = [&](auto &a) {
      a.prm = dprm.p, a.tile_inst = nullptr, a.tile_count = tile_count.p;
      a.epb = til.epb, a.NP = NP, a.tile0 = 0, a.ntiles = ntiles, a.part = part, a.out = out;
    };
    if (prec == MXM_PREC_FP64) {
      SpeedArgs<double> a{};
      a.q = q.p;
      fill(a);
      MXM_HIP(launch_max_speed(phys, true, &a, stream));
    } else {
      SpeedArgs<float> a{};
      a.q = f.p;
      fill(a);
      MXM_HIP(launch_max_speed(phys, false, &a, stream));
    }
  }

  // the largest signal speed over the nodes of the current state
  double max_wave_speed() {
    if (dpart.n < (size_t)ntiles) dpart.alloc(ntiles);
    if (dout.n < 1) dout.alloc(1);
    enqueue_max_speed(dpart.p, dout.p);
    double h = 0.0;
    MXM_HIP(hipMemcpyAsync(&h, dout.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    MXM_HIP(hipStreamSynchronize(stream));
    return h;
  }

  // builds (if needed) the graphs that step(dt, n) replays
  void prepare(double dt, int n) {
    if (n <= 0) return;
    if (dt != g_dt || limiter != g_lim || eps != g_eps || cfg != g_cfg) {
      drop_graphs();
      g_dt = dt, g_lim = limiter, g_eps = eps, g_cfg = cfg;
    }
    if (n >= kGraphSteps && !gfull.exec) capture_steps(dt, kGraphSteps, gfull);
    const int r = n % kGraphSteps;
    if (r && !grem[r].exec) capture_steps(dt, r, grem[r]);
  }

  void step(double dt, int n) {
    if (n <= 0) return;
    prepare(dt, n);
    for (; n >= kGraphSteps; n -= kGraphSteps) MXM_HIP(hipGraphLaunch(gfull.exec, stream));
    if (n) MXM_HIP(hipGraphLaunch(grem[n].exec, stream));
    if (limiter) materialize(0);
  }

  long long dof() const { return (long long)K * NP * NC; }
  size_t state_elem_bytes() const { return prec == MXM_PREC_FP64 ? sizeof(double) : sizeof(float); }
  // q (reference order [K][NP][NC], FP64) <-> device slots and precision
  void upload_state(const double *h) {
    const size_t node = (size_t)NP * NC, n = (size_t)Kd * node;
    MXM_HIP(hipStreamSynchronize(stream));
    if (prec == MXM_PREC_FP64) {
      std::vector<double> d(n, 0.0);
      for (int e = 0; e < K; ++e) std::memcpy(&d[(size_t)til.dev_of[e] * node], h + (size_t)e * node, node * sizeof(double));
      MXM_HIP(hipMemcpy(q.p, d.data(), n * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> d(n, 0.f);
      for (int e = 0; e < K; ++e)
        for (size_t k = 0; k < node; ++k) d[(size_t)til.dev_of[e] * node + k] = float(h[(size_t)e * node + k]);
      MXM_HIP(hipMemcpy(f.p, d.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    MXM_HIP(launch_fill(ltheta.p, (long long)ltheta.n, 1.0, stream));  // nothing pending on a new state
  }
  void download_state(double *h) {
    const size_t node = (size_t)NP * NC, n = (size_t)Kd * node;
    MXM_HIP(hipStreamSynchronize(stream));
    if (prec == MXM_PREC_FP64) {
      std::vector<double> d(n);
      MXM_HIP(hipMemcpy(d.data(), q.p, n * sizeof(double), hipMemcpyDeviceToHost));
      for (int e = 0; e < K; ++e) std::memcpy(h + (size_t)e * node, &d[(size_t)til.dev_of[e] * node], node * sizeof(double));
    } else {
      std::vector<float> d(n);
      MXM_HIP(hipMemcpy(d.data(), f.p, n * sizeof(float), hipMemcpyDeviceToHost));
      for (int e = 0; e < K; ++e)
        for (size_t k = 0; k < node; ++k) h[(size_t)e * node + k] = double(d[(size_t)til.dev_of[e] * node + k]);
    }
  }
  size_t state_bytes() const { return q.bytes() + q1.bytes() + q2.bytes() + f.bytes() + f1.bytes() + f2.bytes(); }
  long long state_bytes_per_step() const { return 8LL * dof() * (long long)state_elem_bytes(); }
};

Solver *as_solver(void *h) {
  if (!h) throw ArgError("null solver handle");
  return static_cast<Solver *>(h);
}

HostMesh *as_mesh(void *h) {
  if (!h) throw ArgError("null mesh handle");
  return static_cast<HostMesh *>(h);
}

int create(int dim, const void *desc, int P, int phys, const double *prm_in, int precision, void **out) {
  if (!out) throw ArgError("null out handle");
  *out = nullptr;
  if (g_device < 0) return fail(MXM_ERR_NOGPU, "mxm_hip_init has not been called");
  if (precision != MXM_PREC_FP64 && precision != MXM_PREC_FP32_MFMA) throw ArgError("unknown precision");
  if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
  const int nc = phys_nc(phys);
  if (nc < 0) throw ArgError("unknown physics id");
  if (dim == 3 && phys > MXM_PHYS_EULER && phys < MXM_PHYS_SW) throw ArgError("2D module physics on a 3D mesh");
  if (dim == 2 && phys >= MXM_PHYS_SW) throw ArgError("3D physics on a 2D mesh");
  auto s = std::make_unique<Solver>();
  s->device = g_device;
  s->phys = phys;
  s->NC = nc;
  s->prec = precision;
  if (prm_in) std::memcpy(s->prm, prm_in, sizeof(s->prm));
  s->variant = phys_variant(phys, s->prm);
  s->cfg = stage_cfg_default(precision);
  s->launch = find_stage(dim, phys, s->variant, P, precision, s->cfg);
  if (!s->launch) throw ArgError("no stage kernel for this (physics, flux, P)");
  RefElem re = build_refelem(dim, P);
  if (dim == 3) {
    auto d = static_cast<const mxm_mesh3d_desc *>(desc);
    check_bc(d->bc, 6);
    s->build(build_mesh_3d(d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->bc, P, re), re);
  } else {
    auto d = static_cast<const mxm_mesh2d_desc *>(desc);
    check_bc(d->bc, 4);
    s->build(build_mesh_2d(d->nx, d->ny, d->lx, d->ly, d->bc, P, re), re);
  }
  *out = s.release();
  return MXM_OK;
}

}  // namespace
}  // namespace mxm

using namespace mxm;

extern "C" {

int mxm_hip_init(int device) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MXM_ERR_NOGPU, "no HIP device");
    if (device < 0 || device >= n) throw ArgError("device index out of range");
    if (device >= kMaxDevices) throw ArgError("device index beyond " + std::to_string(kMaxDevices));
    MXM_HIP(hipSetDevice(device));
    if (!g_scratch[device]) {
      double *p = nullptr;
      MXM_HIP(hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch)));
      if (!p) throw HipError("hipGetSymbolAddress(scratch) returned a null address");
      g_scratch[device] = p;
    }
    g_device = device;
    return MXM_OK;
  });
}

const char *mxm_last_error(void) { return g_err.c_str(); }

int mxm_get_constant(int which, int *value) {
  if (!value) return fail(MXM_ERR_ARG, "null value");
  if (which == 0) {
    *value = MXM_PRM_MAX;
    return MXM_OK;
  }
  if (which == 1) {
    // stage kernels: (dimension, physics, variant, P, form, launch configuration)
    int n = 0;
    for (int prec : {MXM_PREC_FP64, MXM_PREC_FP32_MFMA})
      for (int c = 0; c < stage_cfg_count(prec); ++c)
        for (int P = 1; P <= 5; ++P)
          for (int dim : {2, 3})
            for (int ph = 0; ph <= MXM_PHYS_TWOFLUID; ++ph)
              for (int v = 0; v < 8; ++v) {
                // Euler variants: flux kind + 4 * entropy fix; Euler2D and SW: 2
                const bool valid = ph == MXM_PHYS_EULER ? v != 4 : (ph == 2 || ph == 3) ? v < 2 : v == 0;
                if (valid && find_stage(dim, ph, v, P, prec, c)) ++n;
              }
    *value = n;
    return MXM_OK;
  }
  if (which == 2) {
    *value = stage_cfg_count(MXM_PREC_FP64);
    return MXM_OK;
  }
  if (which == 3) {
    *value = stage_cfg_count(MXM_PREC_FP32_MFMA);
    return MXM_OK;
  }
  return fail(MXM_ERR_ARG, "unknown constant");
}

int mxm_refelem_3d(int P, double *D, double *Lift, double *w, int *f2e) {
  return guarded([&] {
    if (P < 1 || P > 6) throw ArgError("P must be in 1..6");
    RefElem re = build_refelem(3, P);
    if (D) std::memcpy(D, re.D.data(), re.D.size() * sizeof(double));
    if (Lift) std::memcpy(Lift, re.Lift.data(), re.Lift.size() * sizeof(double));
    if (w) std::memcpy(w, re.w.data(), re.w.size() * sizeof(double));
    if (f2e) std::memcpy(f2e, re.fmap.data(), re.fmap.size() * sizeof(int));
    return MXM_OK;
  });
}

int mxm_refelem_2d(int P, double *D, double *Lift, double *w, int *e2e) {
  return guarded([&] {
    if (P < 1 || P > 6) throw ArgError("P must be in 1..6");
    RefElem re = build_refelem(2, P);
    if (D) std::memcpy(D, re.D.data(), re.D.size() * sizeof(double));
    if (Lift) std::memcpy(Lift, re.Lift.data(), re.Lift.size() * sizeof(double));
    if (w) std::memcpy(w, re.w.data(), re.w.size() * sizeof(double));
    if (e2e) std::memcpy(e2e, re.fmap.data(), re.fmap.size() * sizeof(int));
    return MXM_OK;
  });
}

int mxm_refelem_deriv(int dim, int P, double *Dn) {
  return guarded([&] {
    if ((dim != 2 && dim != 3) || P < 1 || P > 6) throw ArgError("dim must be 2 or 3 and P in 1..6");
    if (!Dn) throw ArgError("null output");
    RefElem re = build_refelem(dim, P);
    std::memcpy(Dn, re.Dn.data(), re.Dn.size() * sizeof(double));
    return MXM_OK;
  });
}

int mxm_mesh_create_3d(const mxm_mesh3d_desc *d, int P, void **mesh) {
  return guarded([&] {
    if (!d || !mesh) throw ArgError("null argument");
    *mesh = nullptr;
    if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
    check_bc(d->bc, 6);
    RefElem re = build_refelem(3, P);
    auto m = std::make_unique<HostMesh>(build_mesh_3d(d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->bc, P, re));
    // the kernel-side tables are validated here too, so host-only callers
    // exercise the same checks a solver runs
    build_face_jobs(*m, re, build_connectivity(*m, re), build_tiling(*m, stage_epb(m->dim, m->NP, m->NFP, 1), m->NFP));
    *mesh = m.release();
    return MXM_OK;
  });
}

int mxm_mesh_create_2d(const mxm_mesh2d_desc *d, int P, void **mesh) {
  return guarded([&] {
    if (!d || !mesh) throw ArgError("null argument");
    *mesh = nullptr;
    if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
    check_bc(d->bc, 4);
    RefElem re = build_refelem(2, P);
    auto m = std::make_unique<HostMesh>(build_mesh_2d(d->nx, d->ny, d->lx, d->ly, d->bc, P, re));
    build_face_jobs(*m, re, build_connectivity(*m, re), build_tiling(*m, stage_epb(m->dim, m->NP, m->NFP, 1), m->NFP));
    *mesh = m.release();
    return MXM_OK;
  });
}

int mxm_mesh_info(void *mesh, int *o) {
  return guarded([&] {
    HostMesh *m = as_mesh(mesh);
    if (!o) throw ArgError("null out");
    o[0] = m->dim, o[1] = m->K, o[2] = m->Nf, o[3] = m->NP, o[4] = m->NFP, o[5] = m->NF;
    return MXM_OK;
  });
}

int mxm_mesh_get(void *mesh, int table, void *out) {
  return guarded([&] {
    HostMesh *m = as_mesh(mesh);
    if (!out) throw ArgError("null out");
    auto cpd = [&](const std::vector<double> &v) { std::memcpy(out, v.data(), v.size() * sizeof(double)); };
    auto cpi = [&](const std::vector<int32_t> &v) { std::memcpy(out, v.data(), v.size() * sizeof(int32_t)); };
    switch (table) {
      case MXM_T_ELEM_NODE_XYZ: cpd(m->elem_node_xyz); break;
      case MXM_T_ELEM_INVJ: cpd(m->elem_invJ); break;
      case MXM_T_ELEM_INV_VOL: cpd(m->elem_inv_vol); break;
      case MXM_T_ELEM_FACES: cpi(m->elem_faces); break;
      case MXM_T_ELEM_FACE_SIDE: cpi(m->elem_face_side); break;
      case MXM_T_ELEM_CANON_TO_REF: cpi(m->elem_canon_to_ref); break;
      case MXM_T_FACE_ELEM: cpi(m->face_elem); break;
      case MXM_T_FACE_ELEM_NODE: cpi(m->face_elem_node); break;
      case MXM_T_FACE_NORMAL: cpd(m->face_normal); break;
      case MXM_T_FACE_AREA: cpd(m->face_area); break;
      case MXM_T_FACE_BC_TYPE: cpi(m->face_bc_type); break;
      default: throw ArgError("unknown mesh table");
    }
    return MXM_OK;
  });
}

int mxm_mesh_destroy(void *mesh) {
  delete static_cast<HostMesh *>(mesh);
  return MXM_OK;
}

int mxm_dg_create_3d(const mxm_mesh3d_desc *desc, int P, int phys, const double *prm, int precision,
                     void **solver) {
  return guarded([&] {
    if (!desc) throw ArgError("null descriptor");
    return create(3, desc, P, phys, prm, precision, solver);
  });
}

int mxm_dg_create_2d(const mxm_mesh2d_desc *desc, int P, int phys, const double *prm, int precision,
                     void **solver) {
  return guarded([&] {
    if (!desc) throw ArgError("null descriptor");
    return create(2, desc, P, phys, prm, precision, solver);
  });
}

int mxm_dg_info(void *h, int *o) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!o) throw ArgError("null out");
    o[0] = s->dim, o[1] = s->K, o[2] = s->NP, o[3] = s->NC, o[4] = s->Nf;
    return MXM_OK;
  });
}

int mxm_dg_upload_q(void *h, const double *qh, long long n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!qh || n != s->dof()) throw ArgError("state length must be K*NP*NC");
    s->upload_state(qh);
    return MXM_OK;
  });
}

int mxm_dg_download_q(void *h, double *qh, long long n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!qh || n != s->dof()) throw ArgError("state length must be K*NP*NC");
    s->download_state(qh);
    return MXM_OK;
  });
}

int mxm_dg_download_component(void *h, int c, double *out, long long n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!out || c < 0 || c >= s->NC || n != (long long)s->K * s->NP) throw ArgError("bad component download");
    std::vector<double> all((size_t)s->dof());
    s->download_state(all.data());
    for (long long k = 0; k < n; ++k) out[k] = all[(size_t)k * s->NC + c];
    return MXM_OK;
  });
}

int mxm_dg_rk_stage(void *h, int stage, double dt) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (stage < 0 || stage > 2) throw ArgError("stage must be 0, 1 or 2");
    s->stage(stage, dt);
    if (s->limiter) s->materialize(stage_plan(stage).out);
    return MXM_OK;
  });
}

int mxm_dg_prepare_steps(void *h, double dt, int nsteps) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (nsteps < 0) throw ArgError("negative step count");
    s->prepare(dt, nsteps);
    return MXM_OK;
  });
}

int mxm_dg_step_ssprk3(void *h, double dt, int nsteps) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (nsteps < 0) throw ArgError("negative step count");
    s->step(dt, nsteps);
    return MXM_OK;
  });
}

int mxm_dg_enable_limiter(void *h, int on, double eps) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (s->NC > kLimMaxNC) throw ArgError("limiter supports NC <= " + std::to_string(kLimMaxNC));
    s->limiter = on != 0;
    s->eps = eps;
    return MXM_OK;
  });
}

int mxm_dg_sync(void *h) {
  return guarded([&] {
    MXM_HIP(hipStreamSynchronize(as_solver(h)->stream));
    return MXM_OK;
  });
}

int mxm_dg_bench_step_loop(void *h, double dt, int warmup, int measure, mxm_throughput *out) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!out || warmup < 0 || measure <= 0) throw ArgError("bad bench arguments");
    s->step(dt, warmup);
    MXM_HIP(hipStreamSynchronize(s->stream));
    hipEvent_t ev[2];
    for (auto &e : ev) MXM_HIP(hipEventCreate(&e));
    auto t0 = std::chrono::steady_clock::now();
    MXM_HIP(hipEventRecord(ev[0], s->stream));
    s->step(dt, measure);
    MXM_HIP(hipEventRecord(ev[1], s->stream));
    MXM_HIP(hipStreamSynchronize(s->stream));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0.f;
    const hipError_t erc = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (auto &e : ev) (void)hipEventDestroy(e);
    MXM_HIP(erc);
    const double wall = std::chrono::duration<double>(t1 - t0).count();
    out->steps = measure;
    out->wall_seconds = wall;
    out->dof = s->dof();
    out->state_bytes_per_step = s->state_bytes_per_step();
    out->dof_per_second = (double)s->dof() * measure / wall;
    out->dof_stage_updates_per_second = 3.0 * out->dof_per_second;
    out->state_bandwidth_gbs = (double)s->state_bytes_per_step() * measure / wall * 1e-9;
    const double st = (double)ms * 1e-3 / (3.0 * measure);
    out->stage_seconds = st;
    out->achieved_hbm_gbs = (double)s->state_bytes_per_step() / 3.0 / st * 1e-9;
    out->hbm_fraction = out->achieved_hbm_gbs / kHbmPeakGBs;
    if (s->prec == MXM_PREC_FP32_MFMA) {
      const double flops = 2.0 * (double)s->dof() * (s->dim * s->NP + s->NF * s->NFP);
      out->mfma_tflops = flops / st * 1e-12;
      out->mfma_utilization = out->mfma_tflops / kMfmaF32PeakTFs;
    } else {
      out->mfma_tflops = 0.0;
      out->mfma_utilization = 0.0;
    }
    return MXM_OK;
  });
}

int mxm_dg_dof_count(void *h, long long *out) {
  return guarded([&] {
    if (!out) throw ArgError("null out");
    *out = as_solver(h)->dof();
    return MXM_OK;
  });
}

int mxm_dg_state_bytes_per_step(void *h, long long *out) {
  return guarded([&] {
    if (!out) throw ArgError("null out");
    *out = as_solver(h)->state_bytes_per_step();
    return MXM_OK;
  });
}

int mxm_dg_memory_report(void *h, mxm_memory_report *o) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!o) throw ArgError("null out");
    o->rk_stage = (long long)s->state_bytes();
    o->dg_operators = (long long)(s->Dt.bytes() + s->Lt.bytes() + s->Aop.bytes() + s->w.bytes() + s->fmap.bytes());
    o->limiter = (long long)(s->lmean.bytes() + s->ltheta.bytes() + s->lext.bytes());
    o->mesh_connectivity = (long long)(s->invJ.bytes() + s->inv_vol.bytes() + s->conn.bytes() +
                                       s->jgeo.bytes() + s->jobs.bytes() + s->ipr.bytes() +
                                       s->tile_count.bytes());
    o->halo_device = 0;
    o->halo_pinned = 0;
    return MXM_OK;
  });
}

int mxm_dg_time_stage_kernel(void *h, double dt, int nsteps, double *avg_us) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!avg_us || nsteps <= 0) throw ArgError("bad timing arguments");
    // the limiter is off for the timed steps and restored on every exit
    struct Restore {
      Solver *s;
      bool lim;
      ~Restore() { s->limiter = lim; }
    } restore{s, s->limiter};
    s->limiter = false;
    // two events on the solver's stream around nsteps SSPRK3 steps, launched
    // as the stepping loop launches them (replayed graphs): the average
    // includes the launch-to-launch gaps.  Before the timed region the
    // device is warmed up and the same graphs run once, untimed.
    s->prepare(dt, nsteps);
    s->warm_up(dt, Solver::kWarmUpMs);
    s->step(dt, nsteps);
    hipEvent_t ev[2];
    for (auto &e : ev) MXM_HIP(hipEventCreate(&e));
    MXM_HIP(hipEventRecord(ev[0], s->stream));
    s->step(dt, nsteps);
    MXM_HIP(hipEventRecord(ev[1], s->stream));
    MXM_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    MXM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    *avg_us = (double)ms * 1e3 / (3.0 * nsteps);
    return MXM_OK;
  });
}

int mxm_dg_autotune(void *h, double dt, int reps, int *cfg, double *us, int cap) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!(dt > 0.0) || reps <= 0) throw ArgError("bad autotune arguments");
    double t[kStageCfgMax] = {0};
    const int c = s->autotune(dt, reps, t);
    if (cfg) *cfg = c;
    if (us)
      for (int k = 0; k < cap && k < stage_cfg_count(s->prec); ++k) us[k] = t[k];
    return MXM_OK;
  });
}

int mxm_dg_launch_config(void *h, int cfg, int *current) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (cfg >= 0) {
      StageLaunch l = s->find(cfg, false);
      if (!l) throw ArgError("launch configuration out of range");
      s->cfg = cfg;
      s->launch = l;
    }
    if (current) *current = s->cfg;
    return MXM_OK;
  });
}

int mxm_dg_tile_config(void *h, int wide, int *elems, int *threads) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (wide > 1) throw ArgError("tile size is 0 (256 threads), 1 (wide) or -1 (query)");
    if (wide >= 0) s->set_tile(wide == 1);
    if (elems) *elems = s->til.epb;
    if (threads) *threads = stage_block(s->dim, s->NP, s->NFP, s->NC, s->wide);
    return MXM_OK;
  });
}

int mxm_dg_diagnostics(void *h, double *out, int cap, int *n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    const int nv = 3 * s->NC + 2;
    if (n) *n = nv;
    if (!out || cap < nv) throw ArgError("diagnostics need 3*NC + 2 values");
    s->diagnostics(out);
    return MXM_OK;
  });
}

int mxm_dg_max_wave_speed(void *h, double *speed) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!speed) throw ArgError("null output");
    *speed = s->max_wave_speed();
    return MXM_OK;
  });
}

int mxm_dg_max_wave_speed_noalloc(void *h, double *speed_dev) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!speed_dev) throw ArgError("null output");
    if (s->ntiles > kNoAllocTiles)
      throw ArgError("mesh has " + std::to_string(s->ntiles) + " tiles; the no-allocation reduction holds at most " +
                     std::to_string(kNoAllocTiles));
    double *part = g_scratch[s->device];
    if (!part) return fail(MXM_ERR_NOGPU, "mxm_hip_init has not been called for the solver's device");
    s->enqueue_max_speed(part, speed_dev);
    return MXM_OK;
  });
}

int mxm_dg_destroy(void *h) {
  delete static_cast<Solver *>(h);
  return MXM_OK;
}

}  // extern "C"
