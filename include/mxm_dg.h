/*
 * mxm_dg.h -- C ABI of libmxm_hip.so, the MI355X (gfx950) implementation of
 * mojoxm's DG right-hand side and SSPRK3 time step.
 *
 * Conventions follow the reference's only native shim, src/mpi_shim.c:
 *   - every export is `int mxm_<area>_<verb>(...)`, plain C, no C++ types;
 *   - arguments are int / long long / double, pointers to plain structs
 *     declared here, void* handles and void** out-handles
 *     (src/mpi_shim.c:82-98, src/mpi.mojo:93-97);
 *   - 0 means success, a negative code an error; mxm_last_error() returns
 *     the message of the last failure on the calling thread
 *     (src/mpi.mojo:35-37 raises on rc != 0 the same way);
 *   - one host thread drives a handle (src/mpi_shim.c:41-45, FUNNELED);
 *   - the caller owns host buffers; the library owns all device memory and
 *     its stream.  Uploads/downloads are synchronous; stage and step calls
 *     are asynchronous until a download, mxm_dg_sync or a bench call.
 *
 * State layout (host and device): element-major, q[(e*NP + i)*NC + c]
 * (src/solver.mojo:17-21), FP64 on the host side of the ABI.
 */
#ifndef MXM_DG_H
#define MXM_DG_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define MXM_OK 0
#define MXM_ERR_ARG -1    /* bad argument / unsupported combination        */
#define MXM_ERR_HIP -2    /* a HIP runtime call failed                     */
#define MXM_ERR_NOGPU -3  /* no device / mxm_hip_init not called           */
#define MXM_ERR_MESH -4   /* mesh construction or validation failed        */
#define MXM_ERR_ALLOC -5  /* host or device allocation failed              */

/* ---- boundary kinds, src/boundary.mojo:24-27 --------------------------- */
#define MXM_BC_INTERIOR 0 /* periodic / two-sided face */
#define MXM_BC_WALL 1
#define MXM_BC_OUTFLOW 2
#define MXM_BC_INFLOW 3

/* ---- physics ids and numerical flux kinds ----------------------------- */
#define MXM_PHYS_ADVECTION 0   /* trait Advection, src/advection.mojo, NC=1      */
#define MXM_PHYS_EULER 1       /* trait Euler, src/euler.mojo, NC=5              */
#define MXM_PHYS_EULER2D 2     /* src/local_mesh_2d_gpu_euler.mojo, NC=4 (2D)    */
#define MXM_PHYS_SW2D 3        /* src/local_mesh_2d_gpu_sw.mojo, NC=3 (2D)       */
#define MXM_PHYS_MHDGLM2D 4    /* src/local_mesh_2d_gpu_mhd_glm.mojo, NC=7 (2D)  */
#define MXM_PHYS_MAXWELL2D 5   /* src/local_mesh_2d_gpu_maxwell.mojo, NC=6 (2D)  */
#define MXM_PHYS_ADVECTION2D 6 /* src/local_mesh_2d_gpu_advection.mojo, NC=1 (2D)*/
#define MXM_PHYS_SW 7          /* trait ShallowWater, src/shallow_water.mojo, NC=3 (3D) */
#define MXM_PHYS_MAXWELL 8     /* trait Maxwell, src/maxwell.mojo, NC=6 (3D)     */
#define MXM_PHYS_MHD 9         /* trait IdealMHD + GLM, src/mhd.mojo, NC=9 (3D)  */
#define MXM_PHYS_TWOFLUID 10   /* trait FiveMomentTwoFluid, src/two_fluid.mojo,
                                  NC=17 (3D)                                     */

/* Euler flux kinds, src/euler.mojo:33-36 */
#define MXM_FLUX_RUSANOV 0
#define MXM_FLUX_ROE 1
#define MXM_FLUX_HLLE 2
#define MXM_FLUX_HLLEC 3

/* Physics parameters travel as double prm[MXM_PRM_MAX]; index layout:
 *  ADVECTION   [0..2] v, [3] inflow_q                   (src/advection.mojo:24-40)
 *  EULER       [0] gamma [1] min_density [2] min_pressure [3] flux kind
 *              [4] entropy_fix [5..7] gravity [8..12] inflow (rho, rhou,
 *              rhov, rhow, E)                            (src/euler.mojo:538-575)
 *  EULER2D     [0] gamma [1] min_density [2] min_pressure [3] 0=Rusanov
 *              1=HLLC [4..5] gravity [6..9] inflow (rho, rhou, rhov, E)
 *  SW2D        [0] g [1] min_h [2] 0=Rusanov 1=HLL [3..5] inflow (h, hu, hv)
 *  MHDGLM2D    [0] gamma [1] min_density [2] min_pressure [3] c_h
 *              [4] alpha_d [5..11] inflow (rho, rhou, rhov, Bx, By, E, psi)
 *  MAXWELL2D   [0] c [1..3] J [4..6] M [7..12] inflow (Ex, Ey, Ez, Bx, By, Bz)
 *  ADVECTION2D [0..1] v [2] inflow_q
 *  SW          [0] g [1] h_min [2..4] inflow (h, hu, hv)  (src/shallow_water.mojo:31-52)
 *  MAXWELL     [0] c [1..3] J [4..6] M [7..12] inflow (Ex, Ey, Ez, Bx, By, Bz)
 *                                                         (src/maxwell.mojo:48-110)
 *  MHD         [0] gamma [1] min_density [2] min_pressure [3] c_h [4] alpha_d
 *              [5..13] inflow (rho, rhou, rhov, rhow, E, Bx, By, Bz, psi)
 *                                                         (src/mhd.mojo:197-252)
 *  TWOFLUID    [0] gamma_e [1] gamma_i [2] q_e [3] m_e [4] q_i [5] m_i [6] eps0
 *              [7] c_light [8] c_h [9] alpha_d [10] min_density [11] min_pressure
 *              [12..28] inflow (17 components)            (src/two_fluid.mojo:140-210)
 */
#define MXM_PRM_MAX 32

/* Arithmetic of the stage kernels. */
#define MXM_PREC_FP64 0      /* plain FP64 state, geometry and operators     */
#define MXM_PREC_FP32_MFMA 1 /* FP32 state and fluxes, volume + LIFT on FP32
                                MFMA with FP32 accumulation; geometry,
                                Jacobians, RK combine, positivity rescue and
                                limiter means in FP64                        */

/* ---- mesh descriptors -------------------------------------------------- */
/* bc order [-x, +x, -y, +y, -z, +z], src/boundary.mojo:40-56 */
typedef struct {
    int nx, ny, nz;
    double lx, ly, lz;
    int bc[6];
} mxm_mesh3d_desc;

/* bc order [-x, +x, -y, +y], src/boundary.mojo:64-76 */
typedef struct {
    int nx, ny;
    double lx, ly;
    int bc[4];
} mxm_mesh2d_desc;

/* Mesh tables in the reference's layout (src/local_mesh.mojo:1195-1236,
 * src/local_mesh_2d.mojo:24-40), for mxm_mesh_get. */
#define MXM_T_ELEM_NODE_XYZ 0     /* double [K*NP*ndim]     */
#define MXM_T_ELEM_INVJ 1         /* double [K*ndim*ndim]   */
#define MXM_T_ELEM_INV_VOL 2      /* double [K] 1/(6V), 1/(2A) */
#define MXM_T_ELEM_FACES 3        /* int    [K*NF]          */
#define MXM_T_ELEM_FACE_SIDE 4    /* int    [K*NF]          */
#define MXM_T_ELEM_CANON_TO_REF 5 /* int    [K*NF*NFP]      */
#define MXM_T_FACE_ELEM 6         /* int    [Nf*2]          */
#define MXM_T_FACE_ELEM_NODE 7    /* int    [Nf*2*NFP]      */
#define MXM_T_FACE_NORMAL 8       /* double [Nf*ndim]       */
#define MXM_T_FACE_AREA 9         /* double [Nf]            */
#define MXM_T_FACE_BC_TYPE 10     /* int    [Nf]            */

/* ThroughputReport, src/memory_report.mojo:139-191 */
typedef struct {
    int steps;
    double wall_seconds;
    long long dof;                  /* K*NP*NC, src/solver.mojo:757-763 */
    long long state_bytes_per_step; /* 8 touches * K*NP*NC * sizeof(real) */
    double dof_per_second;          /* dof*steps/wall                   */
    double dof_stage_updates_per_second; /* 3 * dof_per_second          */
    double state_bandwidth_gbs;     /* lower-bound state GB/s           */
    /* Device-side figures (no reference counterpart), from two HIP events
     * on the solver's stream around the measured steps: */
    double stage_seconds;     /* per stage (kernel + limiter when on, gaps included) */
    double achieved_hbm_gbs;  /* state bytes per stage (state_bytes_per_step/3) / stage_seconds */
    double hbm_fraction;      /* achieved_hbm_gbs / 8000 (MI355X HBM3E peak, GB/s) */
    double mfma_tflops;       /* FP32-MFMA form: 2*dof*(ND*NP + NF*NFP) contraction flops per stage / stage_seconds; 0 in FP64 */
    double mfma_utilization;  /* mfma_tflops / 157.3 (MI355X dense FP32 matrix peak); 0 in FP64 */
} mxm_throughput;

/* MemoryReport, src/memory_report.mojo:36-130 (six categories, bytes). */
typedef struct {
    long long rk_stage;
    long long dg_operators;
    long long limiter;
    long long mesh_connectivity;
    long long halo_device;
    long long halo_pinned;
} mxm_memory_report;

/* ---- library ------------------------------------------------------------ */
/* Select the device for this process; replaces mxm_mpi_init's role of
 * process setup (src/mpi_shim.c:41-50). */
int mxm_hip_init(int device);
const char *mxm_last_error(void);
/* Constants by getter, as src/mpi_shim.c:148-160 does for MPI handles:
 * which = 0 -> MXM_PRM_MAX, 1 -> number of stage kernels compiled in,
 * 2 / 3 -> launch configurations of the FP64 / FP32-MFMA stage kernels. */
int mxm_get_constant(int which, int *value);

/* ---- host-only builders (no GPU needed) ---------------------------------- */
/* ReferenceElement[P], src/reference.mojo:519-660: D[3*NP*NP],
 * Lift[4*NP*NFP], w[NP], f2e[4*NFP] (any pointer may be NULL). */
int mxm_refelem_3d(int P, double *D, double *Lift, double *w, int *f2e);
/* ReferenceElement2D[P], src/reference_2d.mojo:130-261: D[2*NP*NP],
 * Lift[3*NP*(P+1)], w[NP], e2e[3*(P+1)]. */
int mxm_refelem_2d(int P, double *D, double *Lift, double *w, int *e2e);
/* Nodal differentiation matrices of the same basis, Dn[dim][NP][NP] with
 * Dn[k][i][j] = d phi_j / d r_k at node i (host only; the diagnostics'
 * vorticity uses them).  The reference has no such matrix: its volume
 * operator is the weak D_ref = M^-1 S (src/reference.mojo:605-620). */
int mxm_refelem_deriv(int dim, int P, double *Dn);
/* LocalMesh[P] (src/local_mesh.mojo:1241-1485) / LocalMesh2D[P]
 * (src/local_mesh_2d.mojo:103-395) built on the host. */
int mxm_mesh_create_3d(const mxm_mesh3d_desc *desc, int P, void **mesh);
int mxm_mesh_create_2d(const mxm_mesh2d_desc *desc, int P, void **mesh);
/* out[0..5] = ndim, K, Nf, NP, NFP, NF */
int mxm_mesh_info(void *mesh, int *out6);
int mxm_mesh_get(void *mesh, int table, void *out);
int mxm_mesh_destroy(void *mesh);

/* ---- solver (replaces Solver[PhysT, P], src/solver.mojo:538-1093, and the
 *      2D <phys>_rk_stage_2d drivers) ------------------------------------- */
int mxm_dg_create_3d(const mxm_mesh3d_desc *desc, int P, int phys, const double *prm,
                     int precision, void **solver);
int mxm_dg_create_2d(const mxm_mesh2d_desc *desc, int P, int phys, const double *prm,
                     int precision, void **solver);
/* out[0..4] = ndim, K, NP, NC, Nf */
int mxm_dg_info(void *solver, int *out5);
/* Copy the whole state d_q from/to the host (n = K*NP*NC doubles). */
int mxm_dg_upload_q(void *solver, const double *q, long long n);
int mxm_dg_download_q(void *solver, double *q, long long n);
/* download_owned_component, src/solver.mojo:711-729: out[K*NP]. */
int mxm_dg_download_component(void *solver, int c, double *out, long long n);
/* One Shu-Osher stage (0, 1, 2) of the plan in src/ssprk3.mojo:41-78,
 * followed by the limiter when it is enabled (src/solver.mojo:1027-1036). */
int mxm_dg_rk_stage(void *solver, int stage, double dt);
/* nsteps x step_ssprk3, src/solver.mojo:1071-1092.  The steps are launched
 * as replayed HIP graphs, one of 32 steps and one of exactly the remainder
 * nsteps mod 32, rebuilt when dt, the limiter or the launch configuration
 * changed since the last call.  With the limiter on, each stage's factors
 * are applied by the next stage as it reads its inputs, and the limited
 * state is written at the end of the call: the state a caller sees is the
 * one per-stage limiting gives, bit for bit. */
int mxm_dg_step_ssprk3(void *solver, double dt, int nsteps);
/* Builds the graphs mxm_dg_step_ssprk3(solver, dt, nsteps) replays, without
 * running a step (no reference counterpart: lets a caller keep graph
 * construction out of a timed region). */
int mxm_dg_prepare_steps(void *solver, double dt, int nsteps);
/* enable_cell_limiter, src/solver.mojo:735-751 (eps: Venkatakrishnan). */
int mxm_dg_enable_limiter(void *solver, int on, double eps);
int mxm_dg_sync(void *solver);
/* bench_step_loop, src/solver.mojo:765-802: warmup steps, sync, measure
 * steps, sync; advances the state by (warmup+measure)*dt. */
int mxm_dg_bench_step_loop(void *solver, double dt, int warmup, int measure,
                           mxm_throughput *out);
/* dof_count / state_bytes_per_step / memory_report,
 * src/solver.mojo:757-914 (FP64 state: 8 bytes per value). */
int mxm_dg_dof_count(void *solver, long long *out);
int mxm_dg_state_bytes_per_step(void *solver, long long *out);
int mxm_dg_memory_report(void *solver, mxm_memory_report *out);
/* Average duration (microseconds) of the stage kernel over 3*nsteps
 * back-to-back launches (limiter off), from two HIP events on the solver's
 * own stream around nsteps SSPRK3 steps launched as mxm_dg_step_ssprk3
 * launches them (replayed HIP graphs), so the average includes the
 * launch-to-launch gaps.  Before them, untimed: about 30 ms of steps (a
 * device coming out of idle clocks up over that much work) and the same
 * nsteps steps once, so the state advances by those steps and 2*nsteps.
 * No reference counterpart: it is the measurement hook that the roofline
 * figures are computed from. */
int mxm_dg_time_stage_kernel(void *solver, double dt, int nsteps, double *avg_us);
/* Domain diagnostics of the current state, computed on the device:
 * out[0..NC) = int q_c dV, out[NC..2NC) = int q_c^2 dV, both mass-weighted
 * (V_e sum_i w_i f(q_i) with the node weights), out[2NC..3NC) = max |q_c|,
 * out[3NC] = kinetic energy int 1/2 rho |u|^2 dV, out[3NC+1] = enstrophy
 * int 1/2 rho |curl u|^2 dV (NaN for physics without a velocity); sums in
 * compensated FP64, in a fixed order.  *n = 3*NC + 2; cap must be >= *n.
 * Replaces DiagnosticsWriter.record's per-component download and host sum
 * (src/diagnostics.mojo:161-223), whose unweighted nodal mean (:170) is
 * replaced by the mass-weighted integral. */
int mxm_dg_diagnostics(void *solver, double *out, int cap, int *n);
/* Largest signal speed over the nodes of the current state, reduced on the
 * device (FP64): |u| + c for the Euler systems, |u| + sqrt(g h) for shallow
 * water, max(|u| + fast speed across B, c_h) for MHD, c for Maxwell, the
 * largest of the two species' |u| + c, c and c_h for the two-fluid system,
 * |v| for advection; NaN if the state holds a NaN.  A caller sets
 * dt = CFL h / (speed (2P + 1)).  No reference counterpart: the
 * reference's drivers use an a-priori speed (examples/euler_vortex_2d_gpu.
 * mojo:148-154) and discard the wave speeds its flux hooks return
 * (src/solver.mojo:229,266,268). */
int mxm_dg_max_wave_speed(void *solver, double *speed);
/* mxm_dg_max_wave_speed without any allocation (no reference counterpart;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * reduction is enqueued on the solver's stream and writes the speed to
 * speed_dev, one double of DEVICE memory owned by the caller; it is there
 * once the stream reaches it (mxm_dg_sync).  The per-tile partials live in
 * a fixed library array of 2^18 doubles per device (2 MiB), so meshes of
 * more than 262144 tiles are rejected with MXM_ERR_ARG.  All calls on one
 * device share that array: calls that overlap on one device race. */
int mxm_dg_max_wave_speed_noalloc(void *solver, double *speed_dev);
/* Launch configurations (no reference counterpart; the reference has one
 * launch shape per kernel).  Each configuration of the stage kernel is an
 * occupancy target (waves per SIMD) that bounds its register allocation;
 * all compute the same bits.  mxm_dg_autotune times each on this solver's
 * mesh -- reps untimed SSPRK3 steps of the current state, then per
 * configuration one untimed step and reps steps between two events; the
 * state is saved to the host first and restored after, so it is left
 * unchanged -- and keeps the fastest: *cfg is its
 * index, us[0..cap) the microseconds per stage of each configuration
 * (mxm_get_constant(2 / 3) = number of configurations of the FP64 /
 * FP32-MFMA form).  mxm_dg_launch_config sets configuration cfg (cfg < 0:
 * no change) and reports the current one.
 * For a shape with a wide tile (the FP32-MFMA form on triangles at P = 3:
 * 32 elements on 320 threads instead of 24 on 256), mxm_dg_autotune also
 * times every configuration on the other tile size and keeps the faster
 * tiling; the state is moved between the two layouts and keeps its values
 * (*cfg and us[] are those of the kept tiling). */
int mxm_dg_autotune(void *solver, double dt, int reps, int *cfg, double *us, int cap);
int mxm_dg_launch_config(void *solver, int cfg, int *current);
/* Tile size of the stage kernel (no reference counterpart): wide = 0 lays
 * the state out in tiles of 256 threads, 1 in wide tiles (MXM_ERR_ARG for
 * a shape or form without them), -1 changes nothing; the state keeps its
 * values.  Reports the current tile's elements and threads per block. */
int mxm_dg_tile_config(void *solver, int wide, int *elems, int *threads);
int mxm_dg_destroy(void *solver);

/* ---- batch engine ----------------------------------------------------------
 * No counterpart in the reference, which advances one problem per Solver /
 * driver loop (the ensemble and serving drivers of examples/ run their
 * problems one after another, e.g. examples/mhd_alfven_glm_2d_gpu.mojo:
 * 215-242).  Many independent problems ("instances", each with its own
 * mesh size, physics parameters, dt and step count) advance together: the
 * instances that share a stage kernel are concatenated and one launch per
 * RK stage steps all of them; one SSPRK3 step is a replayed HIP graph.
 * Every instance follows mxm_dg_step_ssprk3 exactly, and its results do
 * not depend on which other instances are in the batch.  The precision of
 * the batch applies to every instance (MXM_PREC_*). */
int mxm_batch_create(int precision, void **batch);
/* Autotuning of the stage kernels' launch configuration (on by default):
 * when a kernel group is packed, each configuration (see mxm_dg_autotune)
 * is timed on the group's arena and the fastest kept, once per group and
 * power-of-two range of its tile count. */
int mxm_batch_autotune(void *batch, int on);
/* on = 1 (default): each chunk of steps is a replayed HIP graph; on = 0:
 * the same kernels are launched directly on the group's stream (no
 * reference counterpart; for tools that cannot follow graph launches, such
 * as a kernel tracer). */
int mxm_batch_graphs(void *batch, int on);
/* Extra capacity, as a fraction of its tiles and instance slots, of an
 * arena rebuilt because a new instance did not fit (its first build is
 * exact; default 0).  A serving loop that admits work continually sets it
 * so that later arrivals are placed without another rebuild (no reference
 * counterpart). */
int mxm_batch_headroom(void *batch, double fraction);
/* Queue an instance: mesh, order, physics (id + prm[MXM_PRM_MAX]), time
 * step dt, step count, limiter on/off with its eps, initial state q0
 * [K*NP*NC] (reference order, FP64).  It joins at the next
 * mxm_batch_step.  *id names it from then on. */
int mxm_batch_add_2d(void *batch, const mxm_mesh2d_desc *desc, int P, int phys, const double *prm, double dt,
                     int nsteps, int limiter, double eps, const double *q0, long long n, int *id);
int mxm_batch_add_3d(void *batch, const mxm_mesh3d_desc *desc, int P, int phys, const double *prm, double dt,
                     int nsteps, int limiter, double eps, const double *q0, long long n, int *id);
/* Exact device bytes the instance would add to a batch (admission by HBM
 * budget); host only. */
int mxm_batch_footprint_2d(const mxm_mesh2d_desc *desc, int P, int phys, int precision, int limiter,
                           long long *bytes);
int mxm_batch_footprint_3d(const mxm_mesh3d_desc *desc, int P, int phys, int precision, int limiter,
                           long long *bytes);
/* nsteps SSPRK3 steps of every unfinished instance (asynchronous). */
int mxm_batch_step(void *batch, int nsteps);
/* Synchronises; ids[0..*count) = instances that finished all their steps
 * (or failed, mxm_batch_failed) since the last poll (at most cap). */
int mxm_batch_poll(void *batch, int *ids, int cap, int *count);
/* mxm_batch_poll that waits only for the work of the mxm_batch_step calls
 * before the most recent one, so the device stays busy with that one while
 * the host handles the results (no reference counterpart: the serving
 * loop's way of keeping work queued). */
int mxm_batch_poll_lagged(void *batch, int *ids, int cap, int *count);
int mxm_batch_status(void *batch, int id, int *steps_done);
/* *failed = 1 when a stored value of the instance became non-finite: it
 * was switched off at that step (the others go on) and poll reports it as
 * done (no reference counterpart: the serving loop evicts such requests). */
int mxm_batch_failed(void *batch, int id, int *failed);
/* State of one instance, reference order [K*NP*NC], FP64. */
int mxm_batch_get_state(void *batch, int id, double *q, long long n);
/* mxm_dg_diagnostics of one instance. */
int mxm_batch_diagnostics(void *batch, int id, double *out, int cap, int *n);
/* mxm_dg_max_wave_speed of one instance. */
int mxm_batch_max_wave_speed(void *batch, int id, double *speed);
/* Drop an instance.  A placed instance's tiles become a free range that
 * later instances of its group can take without a compaction; its device
 * memory is released when the group is next compacted. */
int mxm_batch_remove(void *batch, int id);
/* Device bytes held by the batch. */
int mxm_batch_memory(void *batch, long long *bytes);
int mxm_batch_sync(void *batch);
int mxm_batch_destroy(void *batch);

#ifdef __cplusplus
}
#endif
#endif
