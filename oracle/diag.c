/*
 * diag.c -- domain diagnostics of a DG state (TEST INFRASTRUCTURE, see
 * oracle.h).
 *
 * The reference's DiagnosticsWriter.record (src/diagnostics.mojo:161-223)
 * reports, per configured component c, int q_c dV and int q_c^2 dV as the
 * Float64 sum of the raw nodal values times domain_volume / total_dof (an
 * unweighted nodal mean, :170), and max |q_c| in Float32.  The build
 * replaces the unweighted mean by the mass-weighted element integral
 *     int_e f dV = V_e * sum_i w_i f(q_i),   w_i = int phi_i / |ref|
 * (node_weights, src/reference.mojo:650-659), which is exact when f is a
 * polynomial of degree <= P over the element, and adds two integrals that
 * the Taylor-Green driver approximates from momentum squares
 * (examples/euler_taylor_green.mojo:181-195):
 *     KE        = int 1/2 rho |u|^2 dV          (nodal: 1/2 |m|^2 / rho)
 *     enstrophy = int 1/2 rho |curl u|^2 dV
 * with u = m / rho at the nodes and the velocity gradient from the nodal
 * differentiation matrices (orc_nodal_deriv) mapped by invJ (rows =
 * reference directions, as in the stage).  Physics with a velocity:
 * EULER and MHD (rho, m[3], ...), EULER2D (rho, m[2], E), MHDGLM2D (rho,
 * m[2], ...), SW2D and SW (h, h u[2]: h in place of rho); the others
 * report NaN for the two.  Sums are carried in long double here (the device form uses
 * compensated FP64 sums), so this is the reference value to ~1e-18.
 *
 * Output layout, NC = the physics' component count:
 *     out[0 .. NC)     int q_c dV
 *     out[NC .. 2NC)   int q_c^2 dV
 *     out[2NC .. 3NC)  max |q_c|
 *     out[3NC]         KE
 *     out[3NC + 1]     enstrophy
 *
 * Pinned by: closed forms on polynomial fields (linear integrals of
 * degree-P fields are exact), the constant state (integral = value x
 * domain volume, zero enstrophy), rigid rotation (curl u = 2 Omega), and
 * the Taylor-Green initial field's KE and enstrophy converging to their
 * closed forms under refinement (tests/test_oracle_diag.py).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

typedef long double real;

/* Velocity component index base and density index of a physics, or -1. */
static int velocity_layout(int phys, int *nv) {
    switch (phys) {
    case ORC_PHYS_EULER:
    case ORC_PHYS_MHD: *nv = 3; return 1;
    case ORC_PHYS_EULER2D:
    case ORC_PHYS_MHDGLM2D:
    case ORC_PHYS_SW2D:
    case ORC_PHYS_SW: *nv = 2; return 1;
    default: *nv = 0; return -1;
    }
}

int orc_diagnostics(const orc_mesh_t *m, const double *w, const double *Dn, int phys, const double *q,
                    double *out) {
    const int NC = orc_phys_nc(phys), NP = m->NP, ND = m->ndim, K = m->num_elements;
    if (NC <= 0) return -1;
    const real ref_vol = ND == 3 ? (real)1 / 6 : (real)1 / 2;
    real *lin = (real *)calloc(NC, sizeof(real));
    real *sq = (real *)calloc(NC, sizeof(real));
    double *mx = (double *)calloc(NC, sizeof(double));
    int nv = 0;
    const int vb = velocity_layout(phys, &nv);
    real ke = 0, ens = 0;
    real *u = (real *)malloc(sizeof(real) * 3 * NP);
    for (int e = 0; e < K; ++e) {
        const real V = ref_vol / (real)m->elem_inv_vol[e];
        const double *qe = q + (size_t)e * NP * NC;
        const double *iJ = &m->elem_invJ[(size_t)e * ND * ND];
        for (int c = 0; c < NC; ++c) {
            real a = 0, b = 0;
            for (int i = 0; i < NP; ++i) {
                const real v = qe[i * NC + c];
                a += w[i] * v;
                b += w[i] * v * v;
                if (fabs(qe[i * NC + c]) > mx[c]) mx[c] = fabs(qe[i * NC + c]);
            }
            lin[c] += V * a;
            sq[c] += V * b;
        }
        if (vb < 0) continue;
        /* u = m / rho; the trait Euler state carries m_z on triangles too */
        for (int i = 0; i < NP; ++i) {
            const real r = qe[i * NC];
            for (int d = 0; d < 3; ++d) u[3 * i + d] = d < nv ? (real)qe[i * NC + vb + d] / r : 0;
        }
        real kee = 0, ense = 0;
        for (int i = 0; i < NP; ++i) {
            const real r = qe[i * NC];
            /* g[a][b] = d u_a / d x_b */
            real g[3][3] = {{0}};
            for (int a = 0; a < 3; ++a)
                for (int k = 0; k < ND; ++k) {
                    real dr = 0;
                    for (int j = 0; j < NP; ++j) dr += (real)Dn[((size_t)k * NP + i) * NP + j] * u[3 * j + a];
                    for (int b = 0; b < ND; ++b) g[a][b] += (real)iJ[k * ND + b] * dr;
                }
            const real w0 = g[2][1] - g[1][2], w1 = g[0][2] - g[2][0], w2 = g[1][0] - g[0][1];
            const real u2 = u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1] + u[3 * i + 2] * u[3 * i + 2];
            kee += w[i] * (real)0.5 * r * u2;
            ense += w[i] * (real)0.5 * r * (w0 * w0 + w1 * w1 + w2 * w2);
        }
        ke += V * kee;
        ens += V * ense;
    }
    for (int c = 0; c < NC; ++c) {
        out[c] = (double)lin[c];
        out[NC + c] = (double)sq[c];
        out[2 * NC + c] = mx[c];
    }
    out[3 * NC] = vb < 0 ? NAN : (double)ke;
    out[3 * NC + 1] = vb < 0 ? NAN : (double)ens;
    free(lin);
    free(sq);
    free(mx);
    free(u);
    return 0;
}
