// solver.hip -- the C ABI of libmxm_hip.so (include/mxm_dg.h): host-side
// builders, the device-resident solver (Solver[PhysT, P] of
// src/solver.mojo:538-1093 and the 2D <phys>_rk_stage_2d drivers) and its
// SSPRK3 stepping, limiter, throughput and memory reports.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mxm_dg.h"
#include "host.hpp"
#include "limiter.hpp"
#include "stage.hpp"

namespace mxm {
namespace {

thread_local std::string g_err;
int g_device = -1;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define MXM_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) throw HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    n = count;
    MXM_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) MXM_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  size_t bytes() const { return n * sizeof(T); }
};

int phys_nc(int phys) {
  static const int nc[7] = {1, 5, 4, 3, 7, 6, 1};
  return (phys >= 0 && phys < 7) ? nc[phys] : -1;
}

int phys_variant(int phys, const double *prm) {
  // Euler: flux kind + 4 * entropy fix (the fix only acts on the wave solvers)
  if (phys == MXM_PHYS_EULER) return (int)prm[3] + (prm[3] != 0.0 && prm[4] != 0.0 ? 4 : 0);
  if (phys == MXM_PHYS_EULER2D) return prm[3] != 0.0 ? 1 : 0;
  if (phys == MXM_PHYS_SW2D) return prm[2] != 0.0 ? 1 : 0;
  return 0;
}

struct Solver {
  int dim = 0, P = 0, NP = 0, NFP = 0, NF = 0, NC = 0, phys = 0;
  int K = 0, Nf = 0, nperm = 0;
  double prm[MXM_PRM_MAX] = {0};
  StageLaunch launch = nullptr;
  hipStream_t stream = nullptr;
  bool limiter = false;
  double eps = 0.1;
  DevBuf<double> q, q1, q2, mean;
  DevBuf<double> invJ, inv_vol, jgeo, Dt, Lt, w;
  DevBuf<int4> conn;
  DevBuf<int> fmap, tile_count;
  DevBuf<int4> jobs;
  DevBuf<unsigned short> ipr;
  int maxj = 0;
  // Device element order: tiles of lattice cells, each contiguous
  // (slot T*epb + el); the host side of the ABI keeps the reference order.
  Tiling til;
  int Kd = 0;  // device element slots, padding included
  int ntiles = 0;

  ~Solver() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  void build(const HostMesh &m, const RefElem &re) {
    dim = m.dim, P = m.P, NP = m.NP, NFP = m.NFP, NF = m.NF, K = m.K, Nf = m.Nf;
    Connectivity cn = build_connectivity(m, re);
    nperm = cn.nperm;
    til = build_tiling(m, stage_epb(dim, NP), NFP);
    Kd = til.slots();
    ntiles = til.ntiles;
    MXM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const size_t nq = (size_t)Kd * NP * NC;
    q.alloc(nq);
    q1.alloc(nq);
    q2.alloc(nq);
    MXM_HIP(hipMemset(q.p, 0, q.bytes()));
    MXM_HIP(hipMemset(q1.p, 0, q1.bytes()));
    MXM_HIP(hipMemset(q2.p, 0, q2.bytes()));
    mean.alloc((size_t)Kd * NC);
    // per-element tables in device order; a padding slot is its own
    // neighbour with a zero state, which the limiter leaves unchanged
    const int dd = dim * dim;
    std::vector<double> ij((size_t)Kd * dd, 0.0), iv((size_t)Kd, 0.0);
    std::vector<int4> cd((size_t)Kd * NF);
    for (int sl = 0; sl < Kd; ++sl) {
      const int e = til.ref_of[sl];
      for (int lf = 0; lf < NF; ++lf) {
        const int32_t *c = &cn.conn[((size_t)(e < 0 ? 0 : e) * NF + lf) * 4];
        cd[(size_t)sl * NF + lf] = e < 0 ? int4{sl, 0, 0, 0} : int4{til.dev_of[c[0]], c[1], c[2], 0};
      }
      if (e < 0) continue;
      for (int k = 0; k < dd; ++k) ij[(size_t)sl * dd + k] = m.elem_invJ[(size_t)e * dd + k];
      iv[sl] = m.elem_inv_vol[e];
    }
    invJ.upload(ij.data(), ij.size());
    inv_vol.upload(iv.data(), iv.size());
    conn.upload(cd.data(), cd.size());
    FaceJobs fj = build_face_jobs(m, re, cn, til);
    maxj = fj.maxj;
    tile_count.upload(til.count.data(), til.count.size());
    jobs.upload(reinterpret_cast<const int4 *>(fj.jobs.data()), fj.jobs.size() / 4);
    jgeo.upload(fj.geo.data(), fj.geo.size());
    ipr.upload(fj.ipr.data(), fj.ipr.size());
    // transposed operators, node columns zero-padded as the kernel expects
    const int npad = stage_node_pad(dim, NP, NC);
    std::vector<double> dt((size_t)dim * NP * npad, 0.0), lt((size_t)NF * NFP * npad, 0.0);
    for (int d = 0; d < dim; ++d)
      for (int i = 0; i < NP; ++i)
        for (int j = 0; j < NP; ++j) dt[((size_t)d * NP + j) * npad + i] = re.D[((size_t)d * NP + i) * NP + j];
    for (int f = 0; f < NF; ++f)
      for (int i = 0; i < NP; ++i)
        for (int r = 0; r < NFP; ++r) lt[((size_t)f * NFP + r) * npad + i] = re.Lift[((size_t)f * NP + i) * NFP + r];
    Dt.upload(dt.data(), dt.size());
    Lt.upload(lt.data(), lt.size());
    w.upload(re.w.data(), re.w.size());
    fmap.upload(re.fmap.data(), re.fmap.size());
    MXM_HIP(hipDeviceSynchronize());
  }

  // Shu-Osher plan of src/ssprk3.mojo:41-78 (FP64 coefficients).
  void stage(int s, double dt) {
    StageArgs<double> a{};
    // (q_in, q_a, q_b, q_out): (q, q, q, q1), (q1, q, q1, q2), (q2, q, q2, q);
    // q_b is always the stage input, which the kernel reads from its tile
    if (s == 0) {
      a.q_in = q.p, a.q_a = q.p, a.q_out = q1.p, a.a = 1.0, a.b = 0.0, a.cdt = 1.0 * dt;
    } else if (s == 1) {
      a.q_in = q1.p, a.q_a = q.p, a.q_out = q2.p, a.a = 0.75, a.b = 0.25, a.cdt = 0.25 * dt;
    } else {
      a.q_in = q2.p, a.q_a = q.p, a.q_out = q.p, a.a = 1.0 / 3.0, a.b = 2.0 / 3.0, a.cdt = (2.0 / 3.0) * dt;
    }
    a.tile_count = tile_count.p;
    a.invJ = invJ.p, a.inv_vol = inv_vol.p;
    a.Dt = Dt.p, a.Lt = Lt.p, a.fmap = fmap.p;
    a.jobs = jobs.p, a.jgeo = jgeo.p, a.ipr = ipr.p, a.maxj = maxj, a.ntiles = ntiles;
    a.post_comp = -1;
    a.post_scale = 1.0;
    // GLM psi damping q[psi] *= exp(-alpha_d dt) once per step after the
    // last stage (src/local_mesh_2d_gpu_mhd_glm.mojo:464-484), fused into
    // the stage-3 epilogue; the limiter is linear per component with theta
    // from component 0, so the two commute.
    if (s == 2 && phys == MXM_PHYS_MHDGLM2D && prm[4] > 0.0) {
      a.post_comp = 6;
      a.post_scale = std::exp(-prm[4] * dt);
    }
    MXM_HIP(launch(a, prm, stream));
    if (limiter) MXM_HIP(launch_limiter(a.q_out, mean.p, w.p, conn.p, Kd, NP, NC, NF, dim, eps, stream));
  }

  void step(double dt, int n) {
    for (int k = 0; k < n; ++k)
      for (int s = 0; s < 3; ++s) stage(s, dt);
  }

  long long dof() const { return (long long)K * NP * NC; }
  // reference order [K][NP][NC] <-> device slots
  void to_device(const double *h, double *d) const {
    const size_t node = (size_t)NP * NC;
    std::fill(d, d + (size_t)Kd * node, 0.0);
    for (int e = 0; e < K; ++e) std::memcpy(d + (size_t)til.dev_of[e] * node, h + (size_t)e * node, node * sizeof(double));
  }
  void to_host(const double *d, double *h) const {
    const size_t node = (size_t)NP * NC;
    for (int e = 0; e < K; ++e) std::memcpy(h + (size_t)e * node, d + (size_t)til.dev_of[e] * node, node * sizeof(double));
  }
  long long state_bytes_per_step() const { return 8LL * dof() * (long long)sizeof(double); }
};

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const HipError &e) {
    return fail(MXM_ERR_HIP, e.what());
  } catch (const ArgError &e) {
    return fail(MXM_ERR_ARG, e.what());
  } catch (const std::bad_alloc &) {
    return fail(MXM_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(MXM_ERR_MESH, e.what());
  }
}

Solver *as_solver(void *h) {
  if (!h) throw ArgError("null solver handle");
  return static_cast<Solver *>(h);
}

HostMesh *as_mesh(void *h) {
  if (!h) throw ArgError("null mesh handle");
  return static_cast<HostMesh *>(h);
}

void check_bc(const int *bc, int n) {
  for (int k = 0; k < n; ++k)
    if (bc[k] < 0 || bc[k] > 3) throw ArgError("boundary kind out of range");
}

int create(int dim, const void *desc, int P, int phys, const double *prm_in, int precision, void **out) {
  if (!out) throw ArgError("null out handle");
  *out = nullptr;
  if (g_device < 0) return fail(MXM_ERR_NOGPU, "mxm_hip_init has not been called");
  if (precision != MXM_PREC_FP64) throw ArgError("only MXM_PREC_FP64 is available");
  if (P < 1 || P > 5) throw ArgError("P must be in 1..5");
  const int nc = phys_nc(phys);
  if (nc < 0) throw ArgError("unknown physics id");
  if (dim == 3 && phys > MXM_PHYS_EULER) throw ArgError("2D module physics on a 3D mesh");
  auto s = std::make_unique<Solver>();
  s->phys = phys;
  s->NC = nc;
  if (prm_in) std::memcpy(s->prm, prm_in, sizeof(s->prm));
  const int variant = phys_variant(phys, s->prm);
  if (dim == 3)
    s->launch = find_stage_3d(phys, variant, P);
  else if (phys <= MXM_PHYS_EULER)
    s->launch = find_stage_2d_trait(phys, variant, P);
  else
    s->launch = find_stage_2d_module(phys, variant, P);
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
    MXM_HIP(hipSetDevice(device));
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
    int n = 0;
    for (int P = 1; P <= 5; ++P) {
      // Euler variants: flux kind + 4 * entropy fix
      for (int v = 0; v < 8; ++v) n += (find_stage_3d(1, v, P) != nullptr) + (find_stage_2d_trait(1, v, P) != nullptr);
      n += (find_stage_3d(0, 0, P) != nullptr) + (find_stage_2d_trait(0, 0, P) != nullptr);
      for (int ph = 2; ph <= 6; ++ph)
        for (int v = 0; v < 2; ++v)
          if ((ph == 2 || ph == 3) || v == 0) n += find_stage_2d_module(ph, v, P) != nullptr;
    }
    *value = n;
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
    build_face_jobs(*m, re, build_connectivity(*m, re), build_tiling(*m, stage_epb(m->dim, m->NP), m->NFP));
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
    build_face_jobs(*m, re, build_connectivity(*m, re), build_tiling(*m, stage_epb(m->dim, m->NP), m->NFP));
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
    std::vector<double> dev(s->q.n);
    s->to_device(qh, dev.data());
    MXM_HIP(hipStreamSynchronize(s->stream));
    MXM_HIP(hipMemcpy(s->q.p, dev.data(), s->q.bytes(), hipMemcpyHostToDevice));
    return MXM_OK;
  });
}

int mxm_dg_download_q(void *h, double *qh, long long n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!qh || n != s->dof()) throw ArgError("state length must be K*NP*NC");
    std::vector<double> dev(s->q.n);
    MXM_HIP(hipStreamSynchronize(s->stream));
    MXM_HIP(hipMemcpy(dev.data(), s->q.p, s->q.bytes(), hipMemcpyDeviceToHost));
    s->to_host(dev.data(), qh);
    return MXM_OK;
  });
}

int mxm_dg_download_component(void *h, int c, double *out, long long n) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (!out || c < 0 || c >= s->NC || n != (long long)s->K * s->NP) throw ArgError("bad component download");
    std::vector<double> dev(s->q.n), all((size_t)s->dof());
    MXM_HIP(hipStreamSynchronize(s->stream));
    MXM_HIP(hipMemcpy(dev.data(), s->q.p, s->q.bytes(), hipMemcpyDeviceToHost));
    s->to_host(dev.data(), all.data());
    for (long long k = 0; k < n; ++k) out[k] = all[(size_t)k * s->NC + c];
    return MXM_OK;
  });
}

int mxm_dg_rk_stage(void *h, int stage, double dt) {
  return guarded([&] {
    Solver *s = as_solver(h);
    if (stage < 0 || stage > 2) throw ArgError("stage must be 0, 1 or 2");
    s->stage(stage, dt);
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
    if (s->NC > kLimMaxNC) throw ArgError("limiter supports NC <= 8");
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
    auto t0 = std::chrono::steady_clock::now();
    s->step(dt, measure);
    MXM_HIP(hipStreamSynchronize(s->stream));
    auto t1 = std::chrono::steady_clock::now();
    const double wall = std::chrono::duration<double>(t1 - t0).count();
    out->steps = measure;
    out->wall_seconds = wall;
    out->dof = s->dof();
    out->state_bytes_per_step = s->state_bytes_per_step();
    out->dof_per_second = (double)s->dof() * measure / wall;
    out->dof_stage_updates_per_second = 3.0 * out->dof_per_second;
    out->state_bandwidth_gbs = (double)s->state_bytes_per_step() * measure / wall * 1e-9;
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
    o->rk_stage = (long long)(s->q.bytes() + s->q1.bytes() + s->q2.bytes());
    o->dg_operators = (long long)(s->Dt.bytes() + s->Lt.bytes() + s->w.bytes() + s->fmap.bytes());
    o->limiter = (long long)s->mean.bytes();
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
    const bool lim = s->limiter;
    s->limiter = false;
    // two events on the solver's stream around 3*nsteps back-to-back stage
    // launches: the average includes the launch-to-launch gaps, as the
    // stepping loop does
    hipEvent_t ev[2];
    for (auto &e : ev) MXM_HIP(hipEventCreate(&e));
    MXM_HIP(hipEventRecord(ev[0], s->stream));
    for (int k = 0; k < nsteps; ++k)
      for (int st = 0; st < 3; ++st) s->stage(st, dt);
    MXM_HIP(hipEventRecord(ev[1], s->stream));
    MXM_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    MXM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    s->limiter = lim;
    *avg_us = (double)ms * 1e3 / (3.0 * nsteps);
    return MXM_OK;
  });
}

int mxm_dg_destroy(void *h) {
  delete static_cast<Solver *>(h);
  return MXM_OK;
}

}  // extern "C"
