/*
 * oracle.h -- CPU FP64 restatement of mojoxm's DG right-hand side and
 * SSPRK3 step.  TEST INFRASTRUCTURE ONLY: the checker that the HIP
 * library is compared against.  Nothing in the shipped library links,
 * loads or calls this code; only tests/, the smoke check and the CPU
 * baseline leg of bench.py do.
 *
 * Every routine restates a reference routine (cited file:line) in double
 * precision with the same formulas and the same evaluation order, so it is
 * the reference algorithm at FP64.  Parity status of each part is recorded
 * in the header comment of the file that implements it.
 */
#ifndef MXM_ORACLE_H
#define MXM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sizes ------------------------------------------------------------ */
int orc_num_tet_nodes(int P); /* (P+1)(P+2)(P+3)/6, src/reference.mojo:60 */
int orc_num_tri_nodes(int P); /* (P+1)(P+2)/2,      src/reference.mojo:64 */

/* ---- reference elements (src/reference.mojo:519-660,
 *      src/reference_2d.mojo:130-261) ------------------------------------ */
/* D[3*NP*NP], Lift[4*NP*NFP], w[NP], f2e[4*NFP], node_pos[NP*3] */
int orc_refelem3d(int P, double *D, double *Lift, double *w, int32_t *f2e,
                  double *node_pos);
/* D[2*NP*NP], Lift[3*NP*(P+1)], w[NP], e2e[3*(P+1)], node_pos[NP*2] */
int orc_refelem2d(int P, double *D, double *Lift, double *w, int32_t *e2e,
                  double *node_pos);
/* Gauss-Jordan inverse, src/reference.mojo:289-331.  Returns -1 if
 * singular. */
int orc_mat_inv(const double *m, int n, double *inv);

/* ---- meshes in the reference's table layout ---------------------------- */
typedef struct {
    int ndim, P, NP, NFP, NF;
    int nx, ny, nz;
    double lx, ly, lz;
    int num_elements, num_faces;
    double *elem_node_xyz;      /* [K*NP*ndim]            */
    double *elem_invJ;          /* [K*ndim*ndim]          */
    double *elem_inv_vol;       /* [K] 1/(6V) or 1/(2A)   */
    int32_t *elem_faces;        /* [K*NF]                 */
    int32_t *elem_face_side;    /* [K*NF]                 */
    int32_t *elem_canon_to_ref; /* [K*NF*NFP]             */
    int32_t *face_elem;         /* [Nf*2]                 */
    int32_t *face_elem_node;    /* [Nf*2*NFP]             */
    double *face_normal;        /* [Nf*ndim]              */
    double *face_area;          /* [Nf] area or length    */
    int32_t *face_bc_type;      /* [Nf]                   */
} orc_mesh_t;

/* bc[6] = [-x,+x,-y,+y,-z,+z] (src/boundary.mojo:40-56); 0 = periodic. */
int orc_mesh3d_counts(int nx, int ny, int nz, const int32_t *bc, int *ne,
                      int *nf);
int orc_mesh3d_build(orc_mesh_t *m, const int32_t *bc);
/* bc[4] = [-x,+x,-y,+y] (src/boundary.mojo:64-76). */
int orc_mesh2d_counts(int nx, int ny, const int32_t *bc, int *ne, int *nf);
int orc_mesh2d_build(orc_mesh_t *m, const int32_t *bc);

/* ---- physics ------------------------------------------------------------ */
enum {
    ORC_PHYS_ADVECTION = 0, /* src/advection.mojo          NC=1 */
    ORC_PHYS_EULER = 1,     /* src/euler.mojo              NC=5 */
    ORC_PHYS_EULER2D = 2,   /* local_mesh_2d_gpu_euler     NC=4 */
    ORC_PHYS_SW2D = 3,      /* local_mesh_2d_gpu_sw        NC=3 */
    ORC_PHYS_MHDGLM2D = 4,  /* local_mesh_2d_gpu_mhd_glm   NC=7 */
    ORC_PHYS_MAXWELL2D = 5, /* local_mesh_2d_gpu_maxwell   NC=6 */
    ORC_PHYS_ADVECTION2D = 6, /* local_mesh_2d_gpu_advection NC=1 */
    ORC_PHYS_SW = 7,         /* src/shallow_water.mojo      NC=3 */
    ORC_PHYS_MAXWELL = 8,    /* src/maxwell.mojo            NC=6 */
    ORC_PHYS_MHD = 9,        /* src/mhd.mojo (GLM)          NC=9 */
    ORC_PHYS_TWOFLUID = 10   /* src/two_fluid.mojo          NC=17 */
};
int orc_phys_nc(int phys);
/* Pointwise hooks of the Physics trait (src/solver.mojo:97-145) for the
 * trait physics (ADVECTION, EULER); used by the flux property tests. */
int orc_internal_flux(int phys, const double *prm, const double *q,
                      double *flux);
int orc_numerical_flux(int phys, const double *prm, const double *ql,
                       const double *qr, double nx, double ny, double nz,
                       double *flux, double *speed);
int orc_boundary_flux(int phys, const double *prm, const double *q, int bc,
                      double nx, double ny, double nz, double *flux,
                      double *speed);
int orc_limit_state(int phys, const double *prm, double *q);
/* The other trait physics (physics3d.c): same hook shapes, flux[d*NC+c]
 * for the internal flux; orc3_source = source_term. */
int orc3_phys_nc(int phys);
int orc3_internal_flux(int phys, const double *prm, const double *q, double *flux);
int orc3_numerical_flux(int phys, const double *prm, const double *ql, const double *qr, double nx, double ny,
                        double nz, double *flux, double *speed);
int orc3_boundary_flux(int phys, const double *prm, const double *q, int bc, double nx, double ny, double nz,
                       double *flux, double *speed);
int orc3_ghost(int phys, const double *prm, const double *q, int bc, double nx, double ny, double nz, double *g);
int orc3_source(int phys, const double *prm, const double *q, double *s);
/* Face flux of a 2D module physics at one face node (ghost state built
 * from bc exactly as the module's face kernel does). */
int orc_face_flux_2d(int phys, const double *prm, const double *ql,
                     const double *qr, int bc, double nx, double ny,
                     double *flux);

/* ---- operators --------------------------------------------------------- */
typedef struct {
    int ndim, P, NP, NFP, NF;
    const double *D;    /* [ndim*NP*NP] */
    const double *Lift; /* [NF*NP*NFP]  */
    const double *w;    /* [NP]         */
} orc_ops_t;

/* One RK stage: q_out = a*q_a + b*q_b + c*dt*L(q_in), then limit_state
 * (3D: src/solver.mojo:161-337; 2D: the module's face-flux kernel followed
 * by its vol_lift_combine_rk kernel).  In 2D the trait physics (ADVECTION,
 * EULER) use nz = 0 and the trait flux hooks. */
int orc_stage(const orc_mesh_t *m, const orc_ops_t *ops, int phys,
              const double *prm, const double *q_in, const double *q_a,
              const double *q_b, double *q_out, double a, double b, double c,
              double dt);
/* Barth-Jespersen limiter on q in place (3D src/solver.mojo:384-530,
 * 2D src/local_mesh_2d_gpu_limiter.mojo:74-153). */
int orc_limiter(const orc_mesh_t *m, const orc_ops_t *ops, int nc, double *q,
                double eps);
/* nsteps SSPRK3 steps at fixed dt (src/solver.mojo:1071-1092,
 * src/ssprk3.mojo:41-78), limiter after every stage when limiter != 0,
 * GLM psi damping after stage 3 for MHDGLM2D. */
int orc_step_ssprk3(const orc_mesh_t *m, const orc_ops_t *ops, int phys,
                    const double *prm, double *q, double dt, int nsteps,
                    int limiter, double eps);
/* Nodal differentiation matrices Dn[k][i][j] = d phi_j / d r_k at node i
 * (ndim*NP*NP). */
int orc_nodal_deriv(int ndim, int P, double *Dn);
/* Domain diagnostics of q[K][NP][NC] (diag.c): out[3*NC + 2] = mass-weighted
 * int q_c dV, int q_c^2 dV, max |q_c|, then KE and enstrophy (NaN for
 * physics without a velocity). */
/* Largest signal speed of one node state over all directions (cfl.c);
 * NaN for an unknown physics or a NaN state. */
double orc_max_speed(int phys, const double *prm, const double *q);

int orc_diagnostics(const orc_mesh_t *m, const double *w, const double *Dn, int phys, const double *q,
                    double *out);
/* OpenMP thread count used by the stage loops (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
