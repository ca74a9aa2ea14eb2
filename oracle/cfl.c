/*
 * cfl.c -- largest signal speed of a node state (TEST INFRASTRUCTURE, see
 * oracle.h).
 *
 * The reference has no CFL reduction: its flux hooks return wave speeds
 * that the solver discards (src/solver.mojo:229,266,268) and its drivers
 * take dt from an a-priori speed (examples/euler_vortex_2d_gpu.mojo:
 * 148-154).  The build's reduction takes, per node, the largest signal
 * speed over all directions, with each physics' density / depth and
 * pressure floors:
 *   advection        |v|
 *   Euler, Euler2D   |u| + sqrt(gamma p / rho)
 *   shallow water    |u| + sqrt(g h)
 *   MHD (2D and 3D)  max(|u| + sqrt((gamma p + |B|^2) / rho), c_h)   (the
 *                    fast speed across B, the largest over directions)
 *   Maxwell          |c|
 *   two-fluid        max over the species of |u_s| + sqrt(gamma_s p_s /
 *                    rho_s), |c|, c_h
 * A state with a NaN component gives NaN.
 * Pinned by the closed forms of tests/test_oracle_cfl.py (states at rest,
 * moving states, the MHD fast speed against the MHD Rusanov flux's own
 * speed for B across n, the largest of the two-fluid signals).
 */
#include <math.h>

#include "oracle.h"

static double floor_at(double v, double f) { return v > f ? v : f; }
static double norm3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

double orc_max_speed(int phys, const double *prm, const double *q) {
    /* a NaN component gives NaN (the floors would otherwise hide a NaN
     * density) */
    const int nc = orc_phys_nc(phys);
    for (int c = 0; c < nc; ++c)
        if (q[c] != q[c]) return NAN;
    switch (phys) {
    case ORC_PHYS_ADVECTION: return norm3(prm[0], prm[1], prm[2]);
    case ORC_PHYS_ADVECTION2D: return norm3(prm[0], prm[1], 0.0);
    case ORC_PHYS_EULER: {
        const double g = prm[0], r = floor_at(q[0], prm[1]);
        const double p = floor_at((g - 1.0) * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / r), prm[2]);
        return norm3(q[1], q[2], q[3]) / r + sqrt(g * p / r);
    }
    case ORC_PHYS_EULER2D: {
        const double g = prm[0], r = floor_at(q[0], prm[1]);
        const double p = floor_at((g - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / r), prm[2]);
        return norm3(q[1], q[2], 0.0) / r + sqrt(g * p / r);
    }
    case ORC_PHYS_SW2D:
    case ORC_PHYS_SW: {
        const double h = floor_at(q[0], prm[1]);
        return norm3(q[1], q[2], 0.0) / h + sqrt(prm[0] * h);
    }
    case ORC_PHYS_MHDGLM2D: {
        const double g = prm[0], r = floor_at(q[0], prm[1]), b2 = q[3] * q[3] + q[4] * q[4];
        const double p = floor_at((g - 1.0) * (q[5] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / r - 0.5 * b2), prm[2]);
        const double s = norm3(q[1], q[2], 0.0) / r + sqrt((g * p + b2) / r);
        return s > prm[3] || s != s ? s : prm[3];
    }
    case ORC_PHYS_MHD: {
        const double g = prm[0], r = floor_at(q[0], prm[1]), b2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
        const double p =
            floor_at((g - 1.0) * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / r - 0.5 * b2), prm[2]);
        const double s = norm3(q[1], q[2], q[3]) / r + sqrt((g * p + b2) / r);
        return s > prm[3] || s != s ? s : prm[3];
    }
    case ORC_PHYS_MAXWELL2D:
    case ORC_PHYS_MAXWELL: return fabs(prm[0]);
    case ORC_PHYS_TWOFLUID: {
        double s = fabs(prm[7]) > prm[8] ? fabs(prm[7]) : prm[8];
        for (int k = 0; k < 2; ++k) {
            const double *qs = q + 5 * k, g = prm[k];
            const double r = floor_at(qs[0], prm[10]);
            const double p = floor_at((g - 1.0) * (qs[4] - 0.5 * (qs[1] * qs[1] + qs[2] * qs[2] + qs[3] * qs[3]) / r),
                                      prm[11]);
            const double v = norm3(qs[1], qs[2], qs[3]) / r + sqrt(g * p / r);
            if (v > s || v != v) s = v;
        }
        return s;
    }
    }
    return NAN;
}
