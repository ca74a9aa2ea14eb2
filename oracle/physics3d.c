/*
 * physics3d.c -- oracle restatement of the other trait physics of the
 * reference's 3D solver (TEST INFRASTRUCTURE, see oracle.h), FP64:
 *
 *   ShallowWater        src/shallow_water.mojo:31-167 (2D system embedded
 *                       in 3D, F^z = 0, Rusanov alpha = |u_n| + sqrt(g h))
 *   Maxwell             src/maxwell.mojo:48-255 (E, B; alpha = c; PEC wall)
 *   IdealMHD (GLM)      src/mhd.mojo:46-367 (9 components; alpha =
 *                       max(|u_n| + c_f, c_h); conducting wall)
 *   FiveMomentTwoFluid  src/two_fluid.mojo:52-513 (electron and ion Euler
 *                       fluids, Maxwell with GLM psi; Lorentz sources)
 *
 * Every face flux is Rusanov, 1/2 (F_n(qL) + F_n(qR)) - 1/2 alpha (qR - qL),
 * with the boundary flux the same formula against the BC ghost state.
 * Parameter layouts (double prm[MXM_PRM_MAX]) are documented in
 * include/mxm_dg.h.  Pinned by tests/test_oracle_physics3d.py:
 * consistency, conservation, rotational invariance, reduction to the
 * Euler Rusanov flux (MHD with B = psi = 0, each two-fluid species with
 * the fields and charges off), the PEC and slip-wall ghost properties, the
 * lake at rest, and the sources in closed form.
 */
#include <math.h>

#include "oracle.h"

static double mx2(double a, double b) { return a > b ? a : b; }
static double ab(double a) { return a >= 0.0 ? a : -a; }

/* generic Rusanov combination */
static void rusanov(int nc, const double *Fl, const double *Fr, const double *ql, const double *qr,
                    double alpha, double *f) {
    for (int c = 0; c < nc; ++c) f[c] = 0.5 * (Fl[c] + Fr[c]) - 0.5 * alpha * (qr[c] - ql[c]);
}

/* ---------------------------------------------------- shallow water -- */
/* normal flux and wave speed of one side, src/shallow_water.mojo:104-121 */
static double sw_side(const double *prm, const double *q, double nx, double ny, double *F) {
    const double g = prm[0], hmin = prm[1];
    const double h = q[0] > hmin ? q[0] : hmin;
    const double u = q[1] / h, v = q[2] / h, un = u * nx + v * ny;
    const double gh2 = 0.5 * g * h * h;
    F[0] = q[1] * nx + q[2] * ny;
    F[1] = q[1] * un + gh2 * nx;
    F[2] = q[2] * un + gh2 * ny;
    return ab(un) + sqrt(g * h);
}

/* ----------------------------------------------------------- Maxwell -- */
/* (F.n)_E = c^2 (B x n), (F.n)_B = n x E, src/maxwell.mojo:151-183 */
static void mw_side(double c2, const double *q, double nx, double ny, double nz, double *F) {
    F[0] = c2 * (q[4] * nz - q[5] * ny);
    F[1] = c2 * (q[5] * nx - q[3] * nz);
    F[2] = c2 * (q[3] * ny - q[4] * nx);
    F[3] = ny * q[2] - nz * q[1];
    F[4] = nz * q[0] - nx * q[2];
    F[5] = nx * q[1] - ny * q[0];
}

/* --------------------------------------------------------------- MHD -- */
static double mhd_rho(const double *q, double rmin) { return q[0] > rmin ? q[0] : rmin; }
/* src/mhd.mojo:50-66 */
static double mhd_p(const double *q, double gamma, double rmin, double pmin) {
    const double rho = mhd_rho(q, rmin);
    const double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / rho;
    const double me = 0.5 * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
    const double p = (gamma - 1.0) * (q[4] - ke - me);
    return p > pmin ? p : pmin;
}
/* fast magnetosonic speed along n, src/mhd.mojo:76-89 */
static double mhd_cf(const double *q, double nx, double ny, double nz, double gamma, double rmin, double pmin) {
    const double rho = mhd_rho(q, rmin), p = mhd_p(q, gamma, rmin, pmin);
    const double bn = q[5] * nx + q[6] * ny + q[7] * nz;
    const double cs2 = gamma * p / rho;
    const double ca2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho;
    const double can2 = bn * bn / rho;
    double disc = (cs2 + ca2) * (cs2 + ca2) - 4.0 * cs2 * can2;
    if (disc < 0.0) disc = 0.0;
    return sqrt(0.5 * (cs2 + ca2 + sqrt(disc)));
}
/* normal flux, src/mhd.mojo:152-190 */
static void mhd_side(const double *prm, const double *q, double nx, double ny, double nz, double *F) {
    const double gamma = prm[0], rmin = prm[1], pmin = prm[2], ch = prm[3];
    const double rho = mhd_rho(q, rmin);
    const double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
    const double E = q[4], bx = q[5], by = q[6], bz = q[7], psi = q[8];
    const double pt = mhd_p(q, gamma, rmin, pmin) + 0.5 * (bx * bx + by * by + bz * bz);
    const double ub = u * bx + v * by + w * bz;
    const double un = u * nx + v * ny + w * nz, bn = bx * nx + by * ny + bz * nz;
    F[0] = rho * un;
    F[1] = rho * u * un + pt * nx - bn * bx;
    F[2] = rho * v * un + pt * ny - bn * by;
    F[3] = rho * w * un + pt * nz - bn * bz;
    F[4] = un * (E + pt) - bn * ub;
    F[5] = un * bx - u * bn + psi * nx;
    F[6] = un * by - v * bn + psi * ny;
    F[7] = un * bz - w * bn + psi * nz;
    F[8] = ch * ch * bn;
}
static double mhd_alpha(const double *prm, const double *q, double nx, double ny, double nz) {
    const double rho = mhd_rho(q, prm[1]);
    const double un = (q[1] * nx + q[2] * ny + q[3] * nz) / rho;
    return ab(un) + mhd_cf(q, nx, ny, nz, prm[0], prm[1], prm[2]);
}

/* --------------------------------------------------------- two-fluid -- */
/* one Euler species, src/two_fluid.mojo:66-125 */
static double sp_rho(const double *q, double rmin) { return q[0] > rmin ? q[0] : rmin; }
static double sp_p(const double *q, double gamma, double rmin, double pmin) {
    const double rho = sp_rho(q, rmin);
    const double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / rho;
    const double p = (gamma - 1.0) * (q[4] - ke);
    return p > pmin ? p : pmin;
}
static double sp_side(const double *q, double gamma, double rmin, double pmin, double nx, double ny, double nz,
                      double *F) {
    const double rho = sp_rho(q, rmin), p = sp_p(q, gamma, rmin, pmin);
    const double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho, un = u * nx + v * ny + w * nz;
    F[0] = rho * un;
    F[1] = rho * u * un + p * nx;
    F[2] = rho * v * un + p * ny;
    F[3] = rho * w * un + p * nz;
    F[4] = un * (q[4] + p);
    return ab(un) + sqrt(gamma * p / rho);
}
/* fields with GLM, src/two_fluid.mojo:301-316 */
static void tf_fields(const double *prm, const double *q, double nx, double ny, double nz, double *F) {
    const double c2 = prm[7] * prm[7], ch2 = prm[8] * prm[8];
    const double *E = q + 10, *B = q + 13, psi = q[16];
    F[10] = c2 * (B[1] * nz - B[2] * ny);
    F[11] = c2 * (B[2] * nx - B[0] * nz);
    F[12] = c2 * (B[0] * ny - B[1] * nx);
    F[13] = (ny * E[2] - nz * E[1]) + psi * nx;
    F[14] = (nz * E[0] - nx * E[2]) + psi * ny;
    F[15] = (nx * E[1] - ny * E[0]) + psi * nz;
    F[16] = ch2 * (B[0] * nx + B[1] * ny + B[2] * nz);
}

/* ------------------------------------------------------- dispatch ---- */
int orc3_phys_nc(int phys) {
    switch (phys) {
    case ORC_PHYS_SW: return 3;
    case ORC_PHYS_MAXWELL: return 6;
    case ORC_PHYS_MHD: return 9;
    case ORC_PHYS_TWOFLUID: return 17;
    }
    return -1;
}

/* flux[d*NC + c] = F^d_c, the reference's internal_flux of each type */
static void sw_internal(const double *prm, const double *q, double *F) {
    /* src/shallow_water.mojo:60-86 */
    const double h = q[0] > prm[1] ? q[0] : prm[1];
    const double hu = q[1], hv = q[2], u = hu / h, v = hv / h, gh2 = 0.5 * prm[0] * h * h;
    F[0] = hu; F[1] = hu * u + gh2; F[2] = hu * v;
    F[3] = hv; F[4] = hu * v; F[5] = hv * v + gh2;
    F[6] = F[7] = F[8] = 0.0;
}
static void mhd_internal(const double *prm, const double *q, double *F) {
    /* mhd_flux_dir, src/mhd.mojo:96-147 */
    const double rho = mhd_rho(q, prm[1]);
    const double mx = q[1], my = q[2], mz = q[3], u = mx / rho, v = my / rho, w = mz / rho;
    const double E = q[4], bx = q[5], by = q[6], bz = q[7], psi = q[8];
    const double pt = mhd_p(q, prm[0], prm[1], prm[2]) + 0.5 * (bx * bx + by * by + bz * bz);
    const double ub = u * bx + v * by + w * bz, ch2 = prm[3] * prm[3];
    double *x = F, *y = F + 9, *z = F + 18;
    x[0] = mx; x[1] = mx * u + pt - bx * bx; x[2] = mx * v - bx * by; x[3] = mx * w - bx * bz;
    x[4] = u * (E + pt) - bx * ub; x[5] = psi; x[6] = u * by - v * bx; x[7] = u * bz - w * bx; x[8] = ch2 * bx;
    y[0] = my; y[1] = my * u - by * bx; y[2] = my * v + pt - by * by; y[3] = my * w - by * bz;
    y[4] = v * (E + pt) - by * ub; y[5] = v * bx - u * by; y[6] = psi; y[7] = v * bz - w * by; y[8] = ch2 * by;
    z[0] = mz; z[1] = mz * u - bz * bx; z[2] = mz * v - bz * by; z[3] = mz * w + pt - bz * bz;
    z[4] = w * (E + pt) - bz * ub; z[5] = w * bx - u * bz; z[6] = w * by - v * bz; z[7] = psi; z[8] = ch2 * bz;
}
static void sp_internal(const double *q, double gamma, double rmin, double pmin, int d, double *F) {
    /* _species_flux_dir, src/two_fluid.mojo:83-110 */
    const double rho = sp_rho(q, rmin), p = sp_p(q, gamma, rmin, pmin);
    const double m[3] = {q[1], q[2], q[3]}, u[3] = {q[1] / rho, q[2] / rho, q[3] / rho};
    F[0] = m[d];
    for (int k = 0; k < 3; ++k) F[1 + k] = m[d] * u[k] + (k == d ? p : 0.0);
    F[4] = u[d] * (q[4] + p);
}

int orc3_internal_flux(int phys, const double *prm, const double *q, double *flux) {
    static const double dirs[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    switch (phys) {
    case ORC_PHYS_SW: sw_internal(prm, q, flux); return 0;
    case ORC_PHYS_MAXWELL:
        /* src/maxwell.mojo:117-148: the normal flux along each axis */
        for (int d = 0; d < 3; ++d) mw_side(prm[0] * prm[0], q, dirs[d][0], dirs[d][1], dirs[d][2], flux + 6 * d);
        return 0;
    case ORC_PHYS_MHD: mhd_internal(prm, q, flux); return 0;
    case ORC_PHYS_TWOFLUID:
        /* src/two_fluid.mojo:213-259 */
        for (int d = 0; d < 3; ++d) {
            double *F = flux + 17 * d;
            sp_internal(q, prm[0], prm[10], prm[11], d, F);
            sp_internal(q + 5, prm[1], prm[10], prm[11], d, F + 5);
            tf_fields(prm, q, dirs[d][0], dirs[d][1], dirs[d][2], F);
        }
        return 0;
    }
    return -1;
}

int orc3_numerical_flux(int phys, const double *prm, const double *ql, const double *qr, double nx, double ny,
                        double nz, double *flux, double *speed) {
    double Fl[17], Fr[17], alpha;
    switch (phys) {
    case ORC_PHYS_SW:
        alpha = mx2(sw_side(prm, ql, nx, ny, Fl), sw_side(prm, qr, nx, ny, Fr));
        rusanov(3, Fl, Fr, ql, qr, alpha, flux);
        break;
    case ORC_PHYS_MAXWELL:
        mw_side(prm[0] * prm[0], ql, nx, ny, nz, Fl);
        mw_side(prm[0] * prm[0], qr, nx, ny, nz, Fr);
        alpha = prm[0];
        rusanov(6, Fl, Fr, ql, qr, alpha, flux);
        break;
    case ORC_PHYS_MHD:
        mhd_side(prm, ql, nx, ny, nz, Fl);
        mhd_side(prm, qr, nx, ny, nz, Fr);
        alpha = mx2(mx2(mhd_alpha(prm, ql, nx, ny, nz), mhd_alpha(prm, qr, nx, ny, nz)), prm[3]);
        rusanov(9, Fl, Fr, ql, qr, alpha, flux);
        break;
    case ORC_PHYS_TWOFLUID: {
        const double ae = mx2(sp_side(ql, prm[0], prm[10], prm[11], nx, ny, nz, Fl),
                              sp_side(qr, prm[0], prm[10], prm[11], nx, ny, nz, Fr));
        const double ai = mx2(sp_side(ql + 5, prm[1], prm[10], prm[11], nx, ny, nz, Fl + 5),
                              sp_side(qr + 5, prm[1], prm[10], prm[11], nx, ny, nz, Fr + 5));
        tf_fields(prm, ql, nx, ny, nz, Fl);
        tf_fields(prm, qr, nx, ny, nz, Fr);
        alpha = mx2(mx2(ae, ai), mx2(prm[7], prm[8]));
        rusanov(17, Fl, Fr, ql, qr, alpha, flux);
        break;
    }
    default: return -1;
    }
    if (speed) *speed = alpha;
    return 0;
}

/* BC ghost states: WALL reflects normal momenta (and for Maxwell-type
 * fields flips tangential E and normal B), INFLOW is the stored state,
 * OUTFLOW copies. */
int orc3_ghost(int phys, const double *prm, const double *q, int bc, double nx, double ny, double nz, double *g) {
    const int nc = orc3_phys_nc(phys);
    if (nc < 0) return -1;
    for (int c = 0; c < nc; ++c) g[c] = q[c];
    const int inflow0 = phys == ORC_PHYS_SW ? 2 : phys == ORC_PHYS_MAXWELL ? 7 : phys == ORC_PHYS_MHD ? 5 : 12;
    if (bc == 3) {
        for (int c = 0; c < nc; ++c) g[c] = prm[inflow0 + c];
        return 0;
    }
    if (bc != 1) return 0;
    double n[3] = {nx, ny, nz};
    /* reflect a momentum-like vector at offset o: v - 2 (v.n) n */
    #define REFLECT(o, nd) do { double vn = 0.0; for (int d = 0; d < nd; ++d) vn += q[(o) + d] * n[d]; \
        for (int d = 0; d < nd; ++d) g[(o) + d] = q[(o) + d] - 2.0 * vn * n[d]; } while (0)
    /* tangential flip of an E-like vector at offset o: 2 (E.n) n - E */
    #define PEC_E(o) do { double en = q[o] * nx + q[(o) + 1] * ny + q[(o) + 2] * nz; \
        for (int d = 0; d < 3; ++d) g[(o) + d] = 2.0 * en * n[d] - q[(o) + d]; } while (0)
    switch (phys) {
    case ORC_PHYS_SW: REFLECT(1, 2); break;                     /* src/shallow_water.mojo:136-142 */
    case ORC_PHYS_MAXWELL: PEC_E(0); REFLECT(3, 3); break;     /* src/maxwell.mojo:203-216 */
    case ORC_PHYS_MHD: REFLECT(1, 3); REFLECT(5, 3); break;    /* src/mhd.mojo:309-320 */
    case ORC_PHYS_TWOFLUID:                                     /* src/two_fluid.mojo:360-381 */
        REFLECT(1, 3); REFLECT(6, 3); PEC_E(10); REFLECT(13, 3); break;
    }
    #undef REFLECT
    #undef PEC_E
    return 0;
}

int orc3_boundary_flux(int phys, const double *prm, const double *q, int bc, double nx, double ny, double nz,
                       double *flux, double *speed) {
    double g[17];
    if (orc3_ghost(phys, prm, q, bc, nx, ny, nz, g)) return -1;
    return orc3_numerical_flux(phys, prm, q, g, nx, ny, nz, flux, speed);
}

int orc3_source(int phys, const double *prm, const double *q, double *s) {
    const int nc = orc3_phys_nc(phys);
    if (nc < 0) return -1;
    for (int c = 0; c < nc; ++c) s[c] = 0.0;
    switch (phys) {
    case ORC_PHYS_MAXWELL: {                                     /* src/maxwell.mojo:245-252 */
        const double c2 = prm[0] * prm[0];
        for (int d = 0; d < 3; ++d) s[d] = -c2 * prm[1 + d];
        for (int d = 0; d < 3; ++d) s[3 + d] = -prm[4 + d];
        break;
    }
    case ORC_PHYS_MHD: s[8] = -prm[4] * q[8]; break;             /* src/mhd.mojo:357-361 */
    case ORC_PHYS_TWOFLUID: {                                    /* src/two_fluid.mojo:388-440 */
        const double re = sp_rho(q, prm[10]), ri = sp_rho(q + 5, prm[10]);
        const double ue[3] = {q[1] / re, q[2] / re, q[3] / re}, ui[3] = {q[6] / ri, q[7] / ri, q[8] / ri};
        const double *E = q + 10, *B = q + 13;
        const double qme = prm[2] / prm[3], qmi = prm[4] / prm[5];
        const double ve[3] = {ue[1] * B[2] - ue[2] * B[1], ue[2] * B[0] - ue[0] * B[2], ue[0] * B[1] - ue[1] * B[0]};
        const double vi[3] = {ui[1] * B[2] - ui[2] * B[1], ui[2] * B[0] - ui[0] * B[2], ui[0] * B[1] - ui[1] * B[0]};
        for (int d = 0; d < 3; ++d) {
            s[1 + d] = qme * re * (E[d] + ve[d]);
            s[6 + d] = qmi * ri * (E[d] + vi[d]);
        }
        s[4] = qme * re * (ue[0] * E[0] + ue[1] * E[1] + ue[2] * E[2]);
        s[9] = qmi * ri * (ui[0] * E[0] + ui[1] * E[1] + ui[2] * E[2]);
        const double ie = 1.0 / prm[6];
        for (int d = 0; d < 3; ++d) s[10 + d] = -(qme * re * ue[d] + qmi * ri * ui[d]) * ie;
        s[16] = -prm[9] * q[16];
        break;
    }
    }
    return 0;
}
