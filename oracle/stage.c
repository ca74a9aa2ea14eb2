/*
 * stage.c -- oracle restatement of one DG right-hand-side + RK stage, the
 * Barth-Jespersen limiters and the SSPRK3 step (TEST INFRASTRUCTURE, see
 * oracle.h), FP64, OpenMP over elements / faces.
 *
 *   3D stage    rk_stage_kernel, src/solver.mojo:161-337 (per element:
 *               phase 1 nodal fluxes, phase 2 one numerical flux per
 *               (local face, canonical slot), phase 3 volume + lift +
 *               source, phase 4 RK combine + limit_state)
 *   2D stage    <phys>_face_flux_kernel_2d then
 *               <phys>_vol_lift_combine_rk_kernel_2d
 *               (src/local_mesh_2d_gpu_{euler,sw,mhd_glm,maxwell,
 *               advection}.mojo); the trait physics (Advection, Euler) on
 *               triangles use the trait hooks with nz = 0 and apply
 *               limit_state after the combine as the 3D Solver does.
 *   limiters    src/solver.mojo:384-530 (3D, theta over all components,
 *               self-neighbour on boundary faces counted);
 *               src/local_mesh_2d_gpu.mojo:246-264 +
 *               src/local_mesh_2d_gpu_limiter.mojo:74-153 (2D, theta from
 *               component 0, boundary faces skipped)
 *   SSPRK3      src/solver.mojo:1071-1092, src/ssprk3.mojo:41-78; GLM psi
 *               damping src/local_mesh_2d_gpu_mhd_glm.mojo:464-484
 *
 * Pinned by tests/test_oracle_stage.py: constant-state preservation for
 * every physics and P (test/euler_2d_gpu_test.mojo:5-18,
 * test/euler_3d_test.mojo:1-16 and the SW / Maxwell / GLM siblings),
 * exact periodic conservation of sum_i w_i q_i, the stage combine and the
 * SSPRK3 step as the cubic Taylor polynomial of the (linear) advection
 * operator, convergence of the DG derivative at rate >= P, closed-form
 * gravity / Maxwell J,M sources and GLM psi damping, the Gaussian return
 * (3D advection), the translated vortex, the PEC cavity TM(1,1) mode, the
 * Sod star-region values of test/sod_exact_riemann_test.mojo:16-19 with the
 * limiter on, and the limiter checks of test/limiter_2d_gpu_test_p3.mojo
 * and test/limiter_3d_test.mojo (passthrough, spike clamp with the far
 * field untouched, cell-mean conservation) plus its scaling form.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

int orc_module_volume_flux(int phys, const double *prm, const double *q, double *Fx, double *Fy);

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#define MAXNC 17
#define MAXNP 64

static int is_trait(int phys) { return phys == ORC_PHYS_ADVECTION || phys == ORC_PHYS_EULER || orc3_phys_nc(phys) > 0; }

/* Trait source_term: Advection none; Euler gravity, src/euler.mojo:738-747;
 * the other trait physics, physics3d.c. */
static void trait_source(int phys, const double *prm, const double *q, double *s) {
    if (phys == ORC_PHYS_EULER) {
        s[0] = 0.0;
        s[1] = q[0] * prm[5];
        s[2] = q[0] * prm[6];
        s[3] = q[0] * prm[7];
        s[4] = q[1] * prm[5] + q[2] * prm[6] + q[3] * prm[7];
    } else if (orc3_phys_nc(phys) > 0) {
        orc3_source(phys, prm, q, s);
    } else {
        s[0] = 0.0;
    }
}

/* ---------------------------------------------------------------- 3D -- */
static int stage3d(const orc_mesh_t *m, const orc_ops_t *ops, int phys, const double *prm,
                   const double *q_in, const double *q_a, const double *q_b, double *q_out, double a,
                   double b, double cc, double dt) {
    const int NP = m->NP, NFP = m->NFP, NC = orc_phys_nc(phys);
    if (!is_trait(phys) || NC < 0) return -1;
    const int K = m->num_elements;
    int nth = orc_get_threads();
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int e = 0; e < K; ++e) {
        double vflux[MAXNP * 3 * MAXNC];
        double fflux[4 * 21 * MAXNC];
        /* phase 1 */
        for (int i = 0; i < NP; ++i) {
            double f[3 * MAXNC];
            orc_internal_flux(phys, prm, q_in + ((size_t)e * NP + i) * NC, f);
            for (int k = 0; k < 3 * NC; ++k) vflux[i * 3 * NC + k] = f[k];
        }
        /* phase 2 */
        for (int fn = 0; fn < 4 * NFP; ++fn) {
            int lf = fn / NFP, mc = fn % NFP;
            int fid = m->elem_faces[e * 4 + lf];
            double nx = m->face_normal[fid * 3], ny = m->face_normal[fid * 3 + 1],
                   nz = m->face_normal[fid * 3 + 2];
            int el = m->face_elem[fid * 2], er = m->face_elem[fid * 2 + 1];
            int nl = m->face_elem_node[fid * 2 * NFP + mc];
            int nr = m->face_elem_node[fid * 2 * NFP + NFP + mc];
            const double *ql = q_in + ((size_t)el * NP + nl) * NC;
            const double *qr = q_in + ((size_t)er * NP + nr) * NC;
            int bc = m->face_bc_type[fid];
            double *fs = &fflux[(lf * NFP + mc) * NC];
            if (bc != 0)
                orc_boundary_flux(phys, prm, ql, bc, nx, ny, nz, fs, NULL);
            else
                orc_numerical_flux(phys, prm, ql, qr, nx, ny, nz, fs, NULL);
        }
        /* phases 3 + 4 */
        const double *iJ = &m->elem_invJ[e * 9];
        double inv6V = m->elem_inv_vol[e];
        for (int i = 0; i < NP; ++i) {
            size_t ob = ((size_t)e * NP + i) * NC;
            double src[MAXNC];
            trait_source(phys, prm, q_in + ob, src);
            for (int c = 0; c < NC; ++c) {
                double vol = 0.0;
                for (int j = 0; j < NP; ++j) {
                    const double *F = &vflux[j * 3 * NC];
                    double fx = F[c], fy = F[NC + c], fz = F[2 * NC + c];
                    double fr0 = iJ[0] * fx + iJ[1] * fy + iJ[2] * fz;
                    double fr1 = iJ[3] * fx + iJ[4] * fy + iJ[5] * fz;
                    double fr2 = iJ[6] * fx + iJ[7] * fy + iJ[8] * fz;
                    double d0 = ops->D[(0 * NP + i) * NP + j];
                    double d1 = ops->D[(1 * NP + i) * NP + j];
                    double d2 = ops->D[(2 * NP + i) * NP + j];
                    vol += fr0 * d0 + fr1 * d1 + fr2 * d2;
                }
                double face = 0.0;
                for (int lf = 0; lf < 4; ++lf) {
                    double sign = m->elem_face_side[e * 4 + lf] == 0 ? 1.0 : -1.0;
                    double area = m->face_area[m->elem_faces[e * 4 + lf]];
                    for (int mc = 0; mc < NFP; ++mc) {
                        int r = m->elem_canon_to_ref[(e * 4 + lf) * NFP + mc];
                        double Lim = ops->Lift[(lf * NP + i) * NFP + r];
                        face += sign * area * Lim * fflux[(lf * NFP + mc) * NC + c];
                    }
                }
                double rhs = vol - inv6V * face + src[c];
                q_out[ob + c] = a * q_a[ob + c] + b * q_b[ob + c] + cc * dt * rhs;
            }
            orc_limit_state(phys, prm, q_out + ob);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- 2D -- */
static int stage2d(const orc_mesh_t *m, const orc_ops_t *ops, int phys, const double *prm,
                   const double *q_in, const double *q_a, const double *q_b, double *q_out, double a,
                   double b, double cc, double dt) {
    const int NP = m->NP, NFP = m->NFP, NC = orc_phys_nc(phys);
    if (NC < 0) return -1;
    const int K = m->num_elements, Nf = m->num_faces;
    int trait = is_trait(phys);
    double *fstar = (double *)malloc(sizeof(double) * (size_t)Nf * NFP * NC);
    int nth = orc_get_threads();
    /* face flux: one value per (face, canonical slot), side 0 is left */
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < Nf * NFP; ++t) {
        int fid = t / NFP, mc = t % NFP;
        double nx = m->face_normal[fid * 2], ny = m->face_normal[fid * 2 + 1];
        int bc = m->face_bc_type[fid];
        int el = m->face_elem[fid * 2], nl = m->face_elem_node[(fid * 2) * NFP + mc];
        const double *ql = q_in + ((size_t)el * NP + nl) * NC;
        int er = m->face_elem[fid * 2 + 1], nr = m->face_elem_node[(fid * 2 + 1) * NFP + mc];
        const double *qr = q_in + ((size_t)er * NP + nr) * NC;
        double *fs = fstar + (size_t)t * NC;
        if (trait) {
            if (bc != 0)
                orc_boundary_flux(phys, prm, ql, bc, nx, ny, 0.0, fs, NULL);
            else
                orc_numerical_flux(phys, prm, ql, qr, nx, ny, 0.0, fs, NULL);
        } else {
            orc_face_flux_2d(phys, prm, ql, qr, bc, nx, ny, fs);
        }
    }
    /* volume + lift + source + RK */
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int e = 0; e < K; ++e) {
        const double *iJ = &m->elem_invJ[e * 4];
        double inv2A = m->elem_inv_vol[e];
        for (int i = 0; i < NP; ++i) {
            double acc[MAXNC] = {0}, face[MAXNC] = {0};
            for (int j = 0; j < NP; ++j) {
                const double *qj = q_in + ((size_t)e * NP + j) * NC;
                double Fx[MAXNC], Fy[MAXNC];
                if (trait) {
                    double F[3 * MAXNC];
                    orc_internal_flux(phys, prm, qj, F);
                    for (int c = 0; c < NC; ++c) {
                        Fx[c] = F[c];
                        Fy[c] = F[NC + c];
                    }
                } else {
                    orc_module_volume_flux(phys, prm, qj, Fx, Fy);
                }
                double Dr = ops->D[(0 * NP + i) * NP + j], Ds = ops->D[(1 * NP + i) * NP + j];
                for (int c = 0; c < NC; ++c)
                    acc[c] += (iJ[0] * Fx[c] + iJ[1] * Fy[c]) * Dr + (iJ[2] * Fx[c] + iJ[3] * Fy[c]) * Ds;
            }
            for (int lf = 0; lf < 3; ++lf) {
                int fid = m->elem_faces[e * 3 + lf];
                double sign = m->elem_face_side[e * 3 + lf] == 0 ? 1.0 : -1.0;
                double flen = m->face_area[fid];
                for (int mc = 0; mc < NFP; ++mc) {
                    int r = m->elem_canon_to_ref[(e * 3 + lf) * NFP + mc];
                    double sLf = sign * flen * ops->Lift[(lf * NP + i) * NFP + r];
                    const double *fs = fstar + ((size_t)fid * NFP + mc) * NC;
                    for (int c = 0; c < NC; ++c) face[c] += sLf * fs[c];
                }
            }
            size_t idx = ((size_t)e * NP + i) * NC;
            const double *qi = q_in + idx;
            double rhs[MAXNC];
            for (int c = 0; c < NC; ++c) rhs[c] = acc[c] - inv2A * face[c];
            if (trait) {
                double src[MAXNC];
                trait_source(phys, prm, qi, src);
                for (int c = 0; c < NC; ++c) rhs[c] = rhs[c] + src[c];
            } else if (phys == ORC_PHYS_EULER2D) {
                /* src/local_mesh_2d_gpu_euler.mojo:126-140 */
                double ir = qi[0] < prm[1] ? prm[1] : qi[0];
                rhs[1] = rhs[1] + ir * prm[4];
                rhs[2] = rhs[2] + ir * prm[5];
                rhs[3] = rhs[3] + (qi[1] * prm[4] + qi[2] * prm[5]);
            } else if (phys == ORC_PHYS_MAXWELL2D) {
                /* src/local_mesh_2d_gpu_maxwell.mojo:133-142 */
                double c2 = prm[0] * prm[0];
                for (int c = 0; c < 3; ++c) rhs[c] = rhs[c] - c2 * prm[1 + c];
                for (int c = 0; c < 3; ++c) rhs[3 + c] = rhs[3 + c] - prm[4 + c];
            }
            for (int c = 0; c < NC; ++c) q_out[idx + c] = a * q_a[idx + c] + b * q_b[idx + c] + cc * dt * rhs[c];
            if (trait) orc_limit_state(phys, prm, q_out + idx);
        }
    }
    free(fstar);
    return 0;
}

int orc_stage(const orc_mesh_t *m, const orc_ops_t *ops, int phys, const double *prm,
              const double *q_in, const double *q_a, const double *q_b, double *q_out, double a,
              double b, double c, double dt) {
    if (m->ndim == 3) return stage3d(m, ops, phys, prm, q_in, q_a, q_b, q_out, a, b, c, dt);
    return stage2d(m, ops, phys, prm, q_in, q_a, q_b, q_out, a, b, c, dt);
}

/* ----------------------------------------------------------- limiter -- */
int orc_limiter(const orc_mesh_t *m, const orc_ops_t *ops, int NC, double *q, double eps) {
    const int K = m->num_elements, NP = m->NP, NF = m->NF;
    double *avg = (double *)malloc(sizeof(double) * (size_t)K * NC);
    double *theta = (double *)malloc(sizeof(double) * (size_t)K);
    int nth = orc_get_threads();
    int three = (m->ndim == 3);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < K * NC; ++t) {
        int e = t / NC, c = t % NC;
        double s = 0.0;
        for (int n = 0; n < NP; ++n) {
            if (three)
                s += ops->w[n] * q[((size_t)e * NP + n) * NC + c];
            else
                s += q[((size_t)e * NP + n) * NC + c] * ops->w[n];
        }
        avg[t] = s;
    }
    double eps2 = eps * eps, tiny = 1.0e-30;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int e = 0; e < K; ++e) {
        int ncl = three ? NC : 1; /* 2D: component 0 only */
        double own[MAXNC], mn[MAXNC], mx[MAXNC], pD[MAXNC], nD[MAXNC];
        for (int c = 0; c < ncl; ++c) own[c] = mn[c] = mx[c] = avg[e * NC + c];
        for (int lf = 0; lf < NF; ++lf) {
            int fid = m->elem_faces[e * NF + lf];
            int el = m->face_elem[fid * 2], er = m->face_elem[fid * 2 + 1];
            int n = (el == e) ? er : el;
            if (!three && n == e) continue;
            for (int c = 0; c < ncl; ++c) {
                double v = avg[n * NC + c];
                if (v < mn[c]) mn[c] = v;
                if (v > mx[c]) mx[c] = v;
            }
        }
        for (int c = 0; c < ncl; ++c) {
            double p = mx[c] - own[c];
            pD[c] = p < 0.0 ? 0.0 : p;
            double n = own[c] - mn[c];
            nD[c] = n < 0.0 ? 0.0 : n;
        }
        double th = 1.0;
        for (int nn = 0; nn < NP; ++nn)
            for (int c = 0; c < ncl; ++c) {
                double d = q[((size_t)e * NP + nn) * NC + c] - own[c];
                double da = d >= 0.0 ? d : -d;
                if (da <= tiny) continue;
                double D = d > 0.0 ? pD[c] : nD[c];
                double D2 = D * D, d2 = da * da, Dd = D * da;
                double al = (D2 + 2.0 * Dd + eps2) / (D2 + 2.0 * d2 + Dd + eps2);
                if (al < th) th = al;
            }
        theta[e] = th;
    }
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int e = 0; e < K; ++e) {
        double th = theta[e];
        if (!(th < 1.0)) continue;
        for (int nn = 0; nn < NP; ++nn)
            for (int c = 0; c < NC; ++c) {
                size_t idx = ((size_t)e * NP + nn) * NC + c;
                double bm = avg[e * NC + c];
                q[idx] = bm + th * (q[idx] - bm);
            }
    }
    free(avg);
    free(theta);
    return 0;
}

/* ------------------------------------------------------------ SSPRK3 -- */
int orc_step_ssprk3(const orc_mesh_t *m, const orc_ops_t *ops, int phys, const double *prm, double *q,
                    double dt, int nsteps, int limiter, double eps) {
    int NC = orc_phys_nc(phys);
    size_t n = (size_t)m->num_elements * m->NP * NC;
    double *q1 = (double *)calloc(n, sizeof(double));
    double *q2 = (double *)calloc(n, sizeof(double));
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        rc |= orc_stage(m, ops, phys, prm, q, q, q, q1, 1.0, 0.0, 1.0, dt);
        if (limiter) orc_limiter(m, ops, NC, q1, eps);
        rc |= orc_stage(m, ops, phys, prm, q1, q, q1, q2, 0.75, 0.25, 0.25, dt);
        if (limiter) orc_limiter(m, ops, NC, q2, eps);
        rc |= orc_stage(m, ops, phys, prm, q2, q, q2, q, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, dt);
        if (limiter) orc_limiter(m, ops, NC, q, eps);
        if (phys == ORC_PHYS_MHDGLM2D && prm[4] > 0.0) {
            double decay = exp(-prm[4] * dt);
            size_t nn = (size_t)m->num_elements * m->NP;
            for (size_t k = 0; k < nn; ++k) q[k * 7 + 6] = q[k * 7 + 6] * decay;
        }
    }
    free(q1);
    free(q2);
    return rc;
}
