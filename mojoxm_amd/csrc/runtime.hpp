// runtime.hpp -- host-side runtime pieces shared by the single solver
// (solver.hip) and the batch engine (batch.hip): error handling of the C
// ABI, device buffers, physics ids, and the per-mesh device tables the
// stage kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mxm_dg.h"
#include "host.hpp"
#include "stage_families.hpp"

namespace mxm {

inline thread_local std::string g_err;
inline int g_device = -1;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define MXM_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) throw ::mxm::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Runs an ABI body, mapping exceptions to the MXM_ERR_* codes and the
// message of mxm_last_error().
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

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    reset();
    n = count;
    MXM_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) MXM_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  size_t bytes() const { return n * sizeof(T); }
};

// A captured step sequence: the graph and its executable live and die
// together (a profiler or the runtime may still refer to the graph's nodes
// while the executable is launched).
struct StepGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

inline int phys_nc(int phys) {
  static const int nc[11] = {1, 5, 4, 3, 7, 6, 1, 3, 6, 9, 17};
  return (phys >= 0 && phys < 11) ? nc[phys] : -1;
}

inline int phys_variant(int phys, const double *prm) {
  // Euler: flux kind + 4 * entropy fix (the fix only acts on the wave solvers)
  if (phys == MXM_PHYS_EULER) return (int)prm[3] + (prm[3] != 0.0 && prm[4] != 0.0 ? 4 : 0);
  if (phys == MXM_PHYS_EULER2D) return prm[3] != 0.0 ? 1 : 0;
  if (phys == MXM_PHYS_SW2D) return prm[2] != 0.0 ? 1 : 0;
  return 0;
}

// launch configurations of each arithmetic form (StageCfg in stage.hpp)
inline int stage_cfg_count(int prec) { return prec == MXM_PREC_FP64 ? StageCfg<double>::count : StageCfg<float>::count; }
inline int stage_cfg_default(int prec) { return prec == MXM_PREC_FP64 ? StageCfg<double>::dflt : StageCfg<float>::dflt; }

// lim: the kernel variant for a limited run (limiter epilogue and the
// inputs' pending factors applied on load); wide: the kernel of the wide
// tile (stage_epb), which reads a state laid out in wide tiles
inline StageLaunch find_stage(int dim, int phys, int variant, int P, int prec, int cfg, bool lim = false,
                              bool wide = false) {
  const bool f64 = prec == MXM_PREC_FP64;
  if (cfg < 0 || cfg >= stage_cfg_count(prec)) return nullptr;
  if (dim == 3) {
    if (phys <= MXM_PHYS_EULER) return (f64 ? stage_3d_fp64 : stage_3d_fp32)(phys, variant, P, cfg, lim, wide);
    return phys >= MXM_PHYS_SW ? (f64 ? stage_3d_sys_fp64 : stage_3d_sys_fp32)(phys, variant, P, cfg, lim, wide)
                               : nullptr;
  }
  if (phys >= MXM_PHYS_SW) return nullptr;
  if (phys <= MXM_PHYS_EULER) return (f64 ? stage_2d_trait_fp64 : stage_2d_trait_fp32)(phys, variant, P, cfg, lim, wide);
  return (f64 ? stage_2d_module_fp64 : stage_2d_module_fp32)(phys, variant, P, cfg, lim, wide);
}

// Times `launch` over `reps` back-to-back launches (after one untimed
// launch) with two events on `stream`; returns microseconds per launch.
inline double time_launches(StageLaunch launch, const void *args, hipStream_t stream, int reps) {
  hipEvent_t ev[2];
  for (auto &e : ev) MXM_HIP(hipEventCreate(&e));
  MXM_HIP(launch(args, stream));
  MXM_HIP(hipEventRecord(ev[0], stream));
  for (int k = 0; k < reps; ++k) MXM_HIP(launch(args, stream));
  MXM_HIP(hipEventRecord(ev[1], stream));
  MXM_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.f;
  MXM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  return (double)ms * 1e3 / reps;
}

// The batch arena's autotuner: times every launch configuration of one
// stage launch (its arguments must leave the launch's inputs unchanged,
// e.g. stage 1 of SSPRK3, which writes q1 from q) and returns the
// fastest; us[k] gets configuration k's time when us is not null.  (A
// solver times whole steps of its state instead: Solver::tune.)
inline int autotune_stage(int dim, int phys, int variant, int P, int prec, const void *args, hipStream_t stream,
                          int reps, double *us, bool lim = false) {
  int best = stage_cfg_default(prec);
  double tbest = 0.0;
  for (int c = 0; c < stage_cfg_count(prec); ++c) {
    StageLaunch l = find_stage(dim, phys, variant, P, prec, c, lim);
    const double t = time_launches(l, args, stream, reps);
    if (us) us[c] = t;
    if (c == 0 || t < tbest) best = c, tbest = t;
  }
  return best;
}

inline void check_bc(const int *bc, int n) {
  for (int k = 0; k < n; ++k)
    if (bc[k] < 0 || bc[k] > 3) throw ArgError("boundary kind out of range");
}

// GLM psi damping q[psi] *= exp(-alpha_d dt) once per step after the last
// stage (src/local_mesh_2d_gpu_mhd_glm.mojo:464-484), fused into the
// stage-3 epilogue; the limiter is linear per component with theta from
// component 0, so the two commute.  Returns the component (-1: none).
inline int post_component(int phys, const double *prm) {
  return phys == MXM_PHYS_MHDGLM2D && prm[4] > 0.0 ? 6 : -1;
}
inline double post_factor(int phys, const double *prm, double dt) {
  return post_component(phys, prm) >= 0 ? std::exp(-prm[4] * dt) : 1.0;
}

// Shu-Osher stage s of src/ssprk3.mojo:41-78 (FP64 coefficients): the
// buffers (q_in, q_a, q_out) as indices into the trio {q, q1, q2} and
// (a, b, c); the b-term buffer is always q_in.
struct StagePlan {
  int in, qa, out;
  double a, b, c;
};
inline StagePlan stage_plan(int s) {
  if (s == 0) return {0, 0, 1, 1.0, 0.0, 1.0};
  if (s == 1) return {1, 0, 2, 0.75, 0.25, 0.25};
  return {2, 0, 0, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0};
}

// Per-mesh tables in device order (slot T*epb + el), local slot numbering.
struct ElemTables {
  Tiling til;
  FaceJobs fj;
  std::vector<double> invJ, inv_vol;
  std::vector<int> conn;  // limiter neighbours: the partner slot of each face (its own slot on a boundary)
};

// wide: tiles of the wide tile size (stage_epb)
inline ElemTables build_elem_tables(const HostMesh &m, const RefElem &re, int nc, bool wide) {
  ElemTables t;
  Connectivity cn = build_connectivity(m, re);
  t.til = build_tiling(m, stage_epb(m.dim, m.NP, m.NFP, nc, wide), m.NFP);
  const int Kd = t.til.slots(), dd = m.dim * m.dim, NF = m.NF;
  t.invJ.assign((size_t)Kd * dd, 0.0);
  t.inv_vol.assign((size_t)Kd, 0.0);
  t.conn.resize((size_t)Kd * NF);
  // a padding slot is its own neighbour with a zero state, which the
  // limiter leaves unchanged
  for (int sl = 0; sl < Kd; ++sl) {
    const int e = t.til.ref_of[sl];
    for (int lf = 0; lf < NF; ++lf) {
      const int32_t *c = &cn.conn[((size_t)(e < 0 ? 0 : e) * NF + lf) * 4];
      t.conn[(size_t)sl * NF + lf] = e < 0 ? sl : t.til.dev_of[c[0]];
    }
    if (e < 0) continue;
    for (int k = 0; k < dd; ++k) t.invJ[(size_t)sl * dd + k] = m.elem_invJ[(size_t)e * dd + k];
    t.inv_vol[sl] = m.elem_inv_vol[e];
  }
  t.fj = build_face_jobs(m, re, cn, t.til);
  return t;
}

// Reference operators in the layout of each arithmetic form.
struct OpTables {
  std::vector<double> Dt, Lt;  // FP64: transposed, node columns zero-padded
  std::vector<float> Aop;      // FP32-MFMA: [D_0 .. D_ND-1 | Lift], 16-row tiles
};

inline OpTables build_operators(const RefElem &re, int dim, int NC, int prec) {
  OpTables o;
  const int NP = re.NP, NFP = re.NFP, NF = re.NF;
  if (prec == MXM_PREC_FP64) {
    const int npad = stage_node_pad(stage_epb(dim, NP, NFP, NC), NP, NC);
    o.Dt.assign((size_t)dim * NP * npad, 0.0);
    o.Lt.assign((size_t)NF * NFP * npad, 0.0);
    for (int d = 0; d < dim; ++d)
      for (int i = 0; i < NP; ++i)
        for (int j = 0; j < NP; ++j) o.Dt[((size_t)d * NP + j) * npad + i] = re.D[((size_t)d * NP + i) * NP + j];
    for (int fc = 0; fc < NF; ++fc)
      for (int i = 0; i < NP; ++i)
        for (int r = 0; r < NFP; ++r)
          o.Lt[((size_t)fc * NFP + r) * npad + i] = re.Lift[((size_t)fc * NP + i) * NFP + r];
  } else {
    // columns: D_0 .. D_dim-1 (dim*NP, zero-padded to kvp), then Lift
    const int mt = mfma_mtiles(NP), kp = mfma_kpad(dim, NP, NFP), kvp = mfma_kvpad(dim, NP);
    o.Aop.assign((size_t)mt * 16 * kp, 0.f);
    for (int i = 0; i < NP; ++i) {
      for (int d = 0; d < dim; ++d)
        for (int j = 0; j < NP; ++j) o.Aop[(size_t)i * kp + d * NP + j] = float(re.D[((size_t)d * NP + i) * NP + j]);
      for (int fc = 0; fc < NF; ++fc)
        for (int r = 0; r < NFP; ++r)
          o.Aop[(size_t)i * kp + kvp + fc * NFP + r] = float(re.Lift[((size_t)fc * NP + i) * NFP + r]);
    }
  }
  return o;
}

}  // namespace mxm
