/*
 * physics.c -- oracle restatement of the pointwise physics hooks
 * (TEST INFRASTRUCTURE, see oracle.h), FP64.
 *
 *   Advection  src/advection.mojo:60-119
 *   Euler      src/euler.mojo:44-792 (Rusanov, Roe, HLLE, HLLEC,
 *              Harten-Hyman fix, BC ghosts, gravity source, limit_state)
 *   2D modules: the face-flux kernels of src/local_mesh_2d_gpu_euler.mojo
 *              (Rusanov 220-337, HLLC 392-545), _sw.mojo (178-426),
 *              _mhd_glm.mojo (236-412), _maxwell.mojo (219-334),
 *              _advection.mojo (249-300), and the nodal fluxes of their
 *              vol_lift_combine_rk kernels.
 *
 * Pinned by tests/test_oracle_physics.py: consistency F*(q,q,n) = F(q).n,
 * conservation F*(qL,qR,n) = -F*(qR,qL,-n), rotational invariance, upwind
 * closed form, Roe exactness on an isolated contact and on a
 * Rankine-Hugoniot shock, HLLC/HLLEC exact contact resolution, wall ghost
 * zero mass flux.  Parameter layouts are the ones documented in
 * include/mxm_dg.h (MXM_PRM_*).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

int orc_phys_nc(int phys) {
    switch (phys) {
    case ORC_PHYS_ADVECTION: return 1;
    case ORC_PHYS_EULER: return 5;
    case ORC_PHYS_EULER2D: return 4;
    case ORC_PHYS_SW2D: return 3;
    case ORC_PHYS_MHDGLM2D: return 7;
    case ORC_PHYS_MAXWELL2D: return 6;
    case ORC_PHYS_ADVECTION2D: return 1;
    }
    return orc3_phys_nc(phys);
}

/* ---------------------------------------------------------- Euler ---- */
typedef struct {
    double gamma, rmin, pmin;
    int flux, efix;
    double g[3], inflow[5];
} euler_t;

static euler_t euler_params(const double *p) {
    euler_t e;
    e.gamma = p[0];
    e.rmin = p[1];
    e.pmin = p[2];
    e.flux = (int)p[3];
    e.efix = p[4] != 0.0;
    for (int d = 0; d < 3; ++d) e.g[d] = p[5 + d];
    for (int c = 0; c < 5; ++c) e.inflow[c] = p[8 + c];
    return e;
}

/* src/euler.mojo:44-56 */
static double rho_fl(const double *q, double rmin) { return q[0] > rmin ? q[0] : rmin; }
static double p_fl(const double *q, double g, double rmin, double pmin) {
    double rho = rho_fl(q, rmin);
    double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / rho;
    double p = (g - 1.0) * (q[4] - ke);
    return p > pmin ? p : pmin;
}

/* src/euler.mojo:64-72 */
static void flux_1d(const euler_t *e, const double *q, double *f) {
    double rho = rho_fl(q, e->rmin), p = p_fl(q, e->gamma, e->rmin, e->pmin);
    double u = q[1] / rho;
    f[0] = q[1];
    f[1] = q[1] * u + p;
    f[2] = q[2] * u;
    f[3] = q[3] * u;
    f[4] = u * (q[4] + p);
}

/* src/euler.mojo:94-118 */
static void construct_tb(double nx, double ny, double nz, double *t, double *b) {
    double ax = nx, ay = ny, az = nz;
    double anx = ax >= 0.0 ? ax : -ax;
    if (anx >= 0.9)
        ay += 1.0;
    else
        ax += 1.0;
    double pro = nx * ax + ny * ay + nz * az;
    ax -= nx * pro;
    ay -= ny * pro;
    az -= nz * pro;
    double mag = sqrt(ax * ax + ay * ay + az * az);
    t[0] = ax / mag;
    t[1] = ay / mag;
    t[2] = az / mag;
    b[0] = ny * t[2] - nz * t[1];
    b[1] = nz * t[0] - nx * t[2];
    b[2] = nx * t[1] - ny * t[0];
}

/* src/euler.mojo:129-143 */
static void rotate(const double *q, const double *n, const double *t, const double *b, double *r) {
    r[0] = q[0];
    r[4] = q[4];
    r[1] = n[0] * q[1] + n[1] * q[2] + n[2] * q[3];
    r[2] = t[0] * q[1] + t[1] * q[2] + t[2] * q[3];
    r[3] = b[0] * q[1] + b[1] * q[2] + b[2] * q[3];
}
static void antirotate(const double *r, const double *n, const double *t, const double *b, double *q) {
    q[0] = r[0];
    q[4] = r[4];
    q[1] = n[0] * r[1] + t[0] * r[2] + b[0] * r[3];
    q[2] = n[1] * r[1] + t[1] * r[2] + b[1] * r[3];
    q[3] = n[2] * r[1] + t[2] * r[2] + b[2] * r[3];
}

typedef struct { double rho, u, v, w, enth, a; } roeavg;

/* src/euler.mojo:161-176 */
static roeavg roe_averages(const euler_t *e, const double *ql, const double *qr, double pl, double pr) {
    roeavg R;
    double g1 = e->gamma - 1.0;
    double rl = rho_fl(ql, e->rmin), rr = rho_fl(qr, e->rmin);
    R.rho = sqrt(rl * rr);
    double sl = sqrt(rl), sr = sqrt(rr), s = sl + sr;
    R.u = (ql[1] / sl + qr[1] / sr) / s;
    R.v = (ql[2] / sl + qr[2] / sr) / s;
    R.w = (ql[3] / sl + qr[3] / sr) / s;
    R.enth = ((ql[4] + pl) / sl + (qr[4] + pr) / sr) / s;
    double u2 = R.u * R.u + R.v * R.v + R.w * R.w;
    double aa2 = dmax(g1 * (R.enth - 0.5 * u2), 0.0);
    R.a = sqrt(aa2);
    return R;
}

/* src/euler.mojo:184-199 */
static void fluctuations(const double *s, const double *wave, double *amdq, double *apdq) {
    for (int m = 0; m < 5; ++m) {
        amdq[m] = 0.0;
        apdq[m] = 0.0;
        for (int mw = 0; mw < 3; ++mw) {
            double sm = s[mw], wm = wave[m * 3 + mw];
            if (sm < 0.0)
                amdq[m] += sm * wm;
            else
                apdq[m] += sm * wm;
        }
    }
}

/* src/euler.mojo:207-288 */
static void harten_hyman(const euler_t *e, const double *ql, const double *qr, const double *s,
                         const double *wave, double *amdq, double *apdq) {
    double g = e->gamma;
    double rho_l = rho_fl(ql, e->rmin), p_l = p_fl(ql, g, e->rmin, e->pmin);
    double c_l = sqrt(g * p_l / rho_l);
    double s0 = ql[1] / rho_l - c_l;
    for (int m = 0; m < 5; ++m) amdq[m] = 0.0;
    int apdq_only = (s0 > 0.0) && (s[0] > 0.0);
    if (!apdq_only) {
        double q10 = dmax(ql[0] + wave[0], e->rmin);
        double q11 = ql[1] + wave[3], q12 = ql[2] + wave[6], q13 = ql[3] + wave[9];
        double q14 = ql[4] + wave[12];
        double ke1 = 0.5 * (q11 * q11 + q12 * q12 + q13 * q13) / q10;
        double p1 = dmax((g - 1.0) * (q14 - ke1), e->pmin);
        double c1 = sqrt(g * p1 / q10);
        double s1 = q11 / q10 - c1;
        double sf;
        if (s0 < 0.0 && s1 > 0.0)
            sf = s0 * (s1 - s[0]) / (s1 - s0);
        else if (s[0] < 0.0)
            sf = s[0];
        else
            sf = 0.0;
        for (int m = 0; m < 5; ++m) amdq[m] = sf * wave[m * 3 + 0];
        if (s[1] < 0.0) {
            for (int m = 0; m < 5; ++m) amdq[m] += s[1] * wave[m * 3 + 1];
            double rho_r = rho_fl(qr, e->rmin), p_r = p_fl(qr, g, e->rmin, e->pmin);
            double c_r = sqrt(g * p_r / rho_r);
            double s3 = qr[1] / rho_r + c_r;
            double q20 = dmax(qr[0] - wave[2], e->rmin);
            double q21 = qr[1] - wave[5], q22 = qr[2] - wave[8], q23 = qr[3] - wave[11];
            double q24 = qr[4] - wave[14];
            double ke2 = 0.5 * (q21 * q21 + q22 * q22 + q23 * q23) / q20;
            double p2 = dmax((g - 1.0) * (q24 - ke2), e->pmin);
            double c2 = sqrt(g * p2 / q20);
            double s2 = q21 / q20 + c2;
            double sf2;
            if (s2 < 0.0 && s3 > 0.0)
                sf2 = s2 * (s3 - s[2]) / (s3 - s2);
            else if (s[2] < 0.0)
                sf2 = s[2];
            else
                sf2 = 0.0;
            for (int m = 0; m < 5; ++m) amdq[m] += sf2 * wave[m * 3 + 2];
        }
    }
    for (int m = 0; m < 5; ++m) {
        double ddf = 0.0;
        for (int mw = 0; mw < 3; ++mw) ddf += s[mw] * wave[m * 3 + mw];
        apdq[m] = ddf - amdq[m];
    }
}

/* src/euler.mojo:296-361 */
static void roe_solver(const euler_t *e, const double *ql, const double *qr, double *s, double *wave,
                       double *amdq, double *apdq) {
    double g1 = e->gamma - 1.0;
    double d0 = qr[0] - ql[0], d1 = qr[1] - ql[1], d2 = qr[2] - ql[2], d3 = qr[3] - ql[3];
    double d4 = qr[4] - ql[4];
    double pl = p_fl(ql, e->gamma, e->rmin, e->pmin), pr = p_fl(qr, e->gamma, e->rmin, e->pmin);
    roeavg ra = roe_averages(e, ql, qr, pl, pr);
    double u = ra.u, v = ra.v, w = ra.w, enth = ra.enth, a = ra.a;
    double u2 = u * u + v * v + w * w;
    double g1a2 = g1 / (a * a);
    double euv = enth - u2;
    double a4 = g1a2 * (euv * d0 + u * d1 + v * d2 + w * d3 - d4);
    double a2 = d2 - v * d0;
    double a3 = d3 - w * d0;
    double a5 = (d1 + (a - u) * d0 - a * a4) / (2.0 * a);
    double a1 = d0 - a4 - a5;
    wave[0] = a1;
    wave[3] = a1 * (u - a);
    wave[6] = a1 * v;
    wave[9] = a1 * w;
    wave[12] = a1 * (enth - u * a);
    s[0] = u - a;
    wave[1] = a4;
    wave[4] = a4 * u;
    wave[7] = a4 * v + a2;
    wave[10] = a4 * w + a3;
    wave[13] = a4 * 0.5 * u2 + a2 * v + a3 * w;
    s[1] = u;
    wave[2] = a5;
    wave[5] = a5 * (u + a);
    wave[8] = a5 * v;
    wave[11] = a5 * w;
    wave[14] = a5 * (enth + u * a);
    s[2] = u + a;
    if (e->efix)
        harten_hyman(e, ql, qr, s, wave, amdq, apdq);
    else
        fluctuations(s, wave, amdq, apdq);
}

/* src/euler.mojo:364-412 */
static void hlle_solver(const euler_t *e, const double *ql, const double *qr, double *s, double *wave,
                        double *amdq, double *apdq) {
    double g = e->gamma;
    double pl = p_fl(ql, g, e->rmin, e->pmin), pr = p_fl(qr, g, e->rmin, e->pmin);
    roeavg ra = roe_averages(e, ql, qr, pl, pr);
    double s_roe_min = ra.u - ra.a, s_roe_max = ra.u + ra.a;
    double rho_l = rho_fl(ql, e->rmin), a_l = sqrt(g * pl / rho_l), u_l = ql[1] / rho_l;
    double s_l_min = u_l - a_l;
    double rho_r = rho_fl(qr, e->rmin), a_r = sqrt(g * pr / rho_r), u_r = qr[1] / rho_r;
    double s_r_max = u_r + a_r;
    s[0] = dmin(s_l_min, s_roe_min);
    s[2] = dmax(s_r_max, s_roe_max);
    s[1] = 0.0;
    double fl[5], fr[5];
    flux_1d(e, ql, fl);
    flux_1d(e, qr, fr);
    double denom = s[0] - s[2];
    for (int m = 0; m < 5; ++m) {
        double qhat = (fr[m] - fl[m] - s[2] * qr[m] + s[0] * ql[m]) / denom;
        wave[m * 3 + 0] = qhat - ql[m];
        wave[m * 3 + 1] = 0.0;
        wave[m * 3 + 2] = qr[m] - qhat;
    }
    if (e->efix)
        harten_hyman(e, ql, qr, s, wave, amdq, apdq);
    else
        fluctuations(s, wave, amdq, apdq);
}

/* src/euler.mojo:415-485 */
static void hllec_solver(const euler_t *e, const double *ql, const double *qr, double *s, double *wave,
                         double *amdq, double *apdq) {
    double g = e->gamma;
    double pl = p_fl(ql, g, e->rmin, e->pmin), pr = p_fl(qr, g, e->rmin, e->pmin);
    roeavg ra = roe_averages(e, ql, qr, pl, pr);
    double s_roe_min = ra.u - ra.a, s_roe_max = ra.u + ra.a;
    double rho_l = rho_fl(ql, e->rmin), a_l = sqrt(g * pl / rho_l), u_l = ql[1] / rho_l;
    double s_l_min = u_l - a_l;
    double rho_r = rho_fl(qr, e->rmin), a_r = sqrt(g * pr / rho_r), u_r = qr[1] / rho_r;
    double s_r_max = u_r + a_r;
    s[0] = dmin(s_l_min, s_roe_min);
    s[2] = dmax(s_r_max, s_roe_max);
    s[1] = (pr - pl + rho_l * u_l * (s[0] - u_l) - rho_r * u_r * (s[2] - u_r)) /
           (rho_l * (s[0] - u_l) - rho_r * (s[2] - u_r));
    double lm = rho_l * (s[0] - u_l) / (s[0] - s[1]);
    double hL[5] = {lm, lm * s[1], lm * ql[2] / rho_l, lm * ql[3] / rho_l,
                    lm * (ql[4] / rho_l + (s[1] - u_l) * (s[1] + pl / (rho_l * (s[0] - u_l))))};
    double rm = rho_r * (s[2] - u_r) / (s[2] - s[1]);
    double hR[5] = {rm, rm * s[1], rm * qr[2] / rho_r, rm * qr[3] / rho_r,
                    rm * (qr[4] / rho_r + (s[1] - u_r) * (s[1] + pr / (rho_r * (s[2] - u_r))))};
    for (int m = 0; m < 5; ++m) {
        wave[m * 3 + 0] = hL[m] - ql[m];
        wave[m * 3 + 1] = hR[m] - hL[m];
        wave[m * 3 + 2] = qr[m] - hR[m];
    }
    if (e->efix)
        harten_hyman(e, ql, qr, s, wave, amdq, apdq);
    else
        fluctuations(s, wave, amdq, apdq);
}

/* src/euler.mojo:495-520 */
static double flux_from_fluct(const euler_t *e, const double *ql, const double *qr, const double *s,
                              const double *amdq, const double *apdq, double *f) {
    double fl[5], fr[5];
    flux_1d(e, ql, fl);
    flux_1d(e, qr, fr);
    for (int m = 0; m < 5; ++m) f[m] = 0.5 * (fl[m] + fr[m]) - 0.5 * (apdq[m] - amdq[m]);
    return dmax(dmax(fabs(s[0]), fabs(s[1])), fabs(s[2]));
}

/* Euler.numerical_flux, src/euler.mojo:612-683 */
static double euler_nflux(const euler_t *e, const double *ql, const double *qr, double nx, double ny,
                          double nz, double *flux) {
    double n[3] = {nx, ny, nz}, t[3], b[3];
    construct_tb(nx, ny, nz, t, b);
    double rl[5], rr[5], rf[5];
    rotate(ql, n, t, b, rl);
    rotate(qr, n, t, b, rr);
    double vmax;
    if (e->flux == 0) {
        double rhol = rho_fl(rl, e->rmin), rhor = rho_fl(rr, e->rmin);
        double pl = p_fl(rl, e->gamma, e->rmin, e->pmin), pr = p_fl(rr, e->gamma, e->rmin, e->pmin);
        double cl = sqrt(e->gamma * pl / rhol), cr = sqrt(e->gamma * pr / rhor);
        double uxl = rl[1] / rhol, uxr = rr[1] / rhor;
        double cmax = dmax(fabs(uxl) + cl, fabs(uxr) + cr);
        rf[0] = 0.5 * (rl[1] + rr[1] + cmax * (rhol - rhor));
        rf[1] = 0.5 * (rl[1] * uxl + pl + rr[1] * uxr + pr + cmax * (rl[1] - rr[1]));
        rf[2] = 0.5 * (rl[2] * uxl + rr[2] * uxr + cmax * (rl[2] - rr[2]));
        rf[3] = 0.5 * (rl[3] * uxl + rr[3] * uxr + cmax * (rl[3] - rr[3]));
        rf[4] = 0.5 * ((rl[4] + pl) * uxl + (rr[4] + pr) * uxr + cmax * (rl[4] - rr[4]));
        vmax = cmax;
    } else {
        double s[3], wave[15], amdq[5], apdq[5];
        if (e->flux == 1)
            roe_solver(e, rl, rr, s, wave, amdq, apdq);
        else if (e->flux == 2)
            hlle_solver(e, rl, rr, s, wave, amdq, apdq);
        else
            hllec_solver(e, rl, rr, s, wave, amdq, apdq);
        vmax = flux_from_fluct(e, rl, rr, s, amdq, apdq, rf);
    }
    antirotate(rf, n, t, b, flux);
    return vmax;
}

/* ------------------------------------------------------ trait hooks -- */
int orc_internal_flux(int phys, const double *prm, const double *q, double *flux) {
    if (phys == ORC_PHYS_ADVECTION) {
        flux[0] = prm[0] * q[0];
        flux[1] = prm[1] * q[0];
        flux[2] = prm[2] * q[0];
        return 0;
    }
    if (phys == ORC_PHYS_EULER) {
        euler_t e = euler_params(prm);
        double rho = rho_fl(q, e.rmin), p = p_fl(q, e.gamma, e.rmin, e.pmin);
        double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
        flux[0] = q[1]; flux[1] = q[1] * u + p; flux[2] = q[1] * v; flux[3] = q[1] * w;
        flux[4] = u * (q[4] + p);
        flux[5] = q[2]; flux[6] = q[2] * u; flux[7] = q[2] * v + p; flux[8] = q[2] * w;
        flux[9] = v * (q[4] + p);
        flux[10] = q[3]; flux[11] = q[3] * u; flux[12] = q[3] * v; flux[13] = q[3] * w + p;
        flux[14] = w * (q[4] + p);
        return 0;
    }
    return orc3_internal_flux(phys, prm, q, flux);
}

int orc_numerical_flux(int phys, const double *prm, const double *ql, const double *qr, double nx,
                       double ny, double nz, double *flux, double *speed) {
    double sp;
    if (phys == ORC_PHYS_ADVECTION) {
        double nc = prm[0] * nx + prm[1] * ny + prm[2] * nz;
        double an = nc >= 0.0 ? nc : -nc;
        flux[0] = 0.5 * ((nc + an) * ql[0] + (nc - an) * qr[0]);
        sp = an;
    } else if (phys == ORC_PHYS_EULER) {
        euler_t e = euler_params(prm);
        sp = euler_nflux(&e, ql, qr, nx, ny, nz, flux);
    } else {
        return orc3_numerical_flux(phys, prm, ql, qr, nx, ny, nz, flux, speed);
    }
    if (speed) *speed = sp;
    return 0;
}

int orc_boundary_flux(int phys, const double *prm, const double *q, int bc, double nx, double ny,
                      double nz, double *flux, double *speed) {
    if (phys == ORC_PHYS_ADVECTION) {
        /* src/advection.mojo:98-112 */
        double nc = prm[0] * nx + prm[1] * ny + prm[2] * nz;
        double an = nc >= 0.0 ? nc : -nc;
        double qg = (bc == 3) ? prm[3] : 0.0;
        flux[0] = 0.5 * ((nc + an) * q[0] + (nc - an) * qg);
        if (speed) *speed = an;
        return 0;
    }
    if (phys == ORC_PHYS_EULER) {
        /* src/euler.mojo:705-730 */
        euler_t e = euler_params(prm);
        double g[5];
        if (bc == 1) {
            g[0] = q[0];
            g[4] = q[4];
            double mn = q[1] * nx + q[2] * ny + q[3] * nz;
            g[1] = q[1] - 2.0 * mn * nx;
            g[2] = q[2] - 2.0 * mn * ny;
            g[3] = q[3] - 2.0 * mn * nz;
        } else if (bc == 3) {
            for (int c = 0; c < 5; ++c) g[c] = e.inflow[c];
        } else {
            for (int c = 0; c < 5; ++c) g[c] = q[c];
        }
        double sp = euler_nflux(&e, q, g, nx, ny, nz, flux);
        if (speed) *speed = sp;
        return 0;
    }
    return orc3_boundary_flux(phys, prm, q, bc, nx, ny, nz, flux, speed);
}

/* Euler.limit_state, src/euler.mojo:758-792 */
int orc_limit_state(int phys, const double *prm, double *q) {
    if (phys != ORC_PHYS_EULER) return 0;
    euler_t e = euler_params(prm);
    double rho = q[0];
    if (!(rho > e.rmin)) {
        q[0] = e.rmin;
        q[1] = q[2] = q[3] = 0.0;
        q[4] = e.pmin / (e.gamma - 1.0);
        return 0;
    }
    double m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (!(m2 >= 0.0)) {
        q[1] = q[2] = q[3] = 0.0;
        q[4] = e.pmin / (e.gamma - 1.0);
        return 0;
    }
    double ke = 0.5 * m2 / rho;
    double p = (e.gamma - 1.0) * (q[4] - ke);
    if (!(p > e.pmin)) q[4] = ke + e.pmin / (e.gamma - 1.0);
    return 0;
}

/* --------------------------------------------------- 2D module fluxes -- */

/* Euler2D NC=4 nodal flux + Rusanov speed (src/local_mesh_2d_gpu_euler.mojo:283-329) */
static void e2d_state(const double *prm, const double *q, double *Fx, double *Fy, double *speed,
                      double *p_out, double *rho_out) {
    double g = prm[0], rmin = prm[1], pmin = prm[2];
    double rho = q[0];
    if (rho < rmin) rho = rmin;
    double u = q[1] / rho, v = q[2] / rho;
    double ke = 0.5 * rho * (u * u + v * v);
    double p = (g - 1.0) * (q[3] - ke);
    if (p < pmin) p = pmin;
    Fx[0] = q[1]; Fx[1] = q[1] * u + p; Fx[2] = q[1] * v; Fx[3] = u * (q[3] + p);
    Fy[0] = q[2]; Fy[1] = q[2] * u; Fy[2] = q[2] * v + p; Fy[3] = v * (q[3] + p);
    if (speed) *speed = sqrt(u * u + v * v) + sqrt(g * p / rho);
    if (p_out) *p_out = p;
    if (rho_out) *rho_out = rho;
}

static void e2d_hllc(const double *prm, const double *qL, const double *qR, double nx, double ny,
                     double *f) {
    /* src/local_mesh_2d_gpu_euler.mojo:466-545 */
    double g = prm[0], rmin = prm[1], pmin = prm[2];
    double rho_L = qL[0] < rmin ? rmin : qL[0];
    double rho_R = qR[0] < rmin ? rmin : qR[0];
    double uL = qL[1] / rho_L, vL = qL[2] / rho_L, uR = qR[1] / rho_R, vR = qR[2] / rho_R;
    double keL = 0.5 * rho_L * (uL * uL + vL * vL), keR = 0.5 * rho_R * (uR * uR + vR * vR);
    double pL = (g - 1.0) * (qL[3] - keL);
    if (pL < pmin) pL = pmin;
    double pR = (g - 1.0) * (qR[3] - keR);
    if (pR < pmin) pR = pmin;
    double aL = sqrt(g * pL / rho_L), aR = sqrt(g * pR / rho_R);
    double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
    double FnL[4] = {qL[1] * nx + qL[2] * ny, (qL[1] * uL + pL) * nx + (qL[1] * vL) * ny,
                     (qL[1] * vL) * nx + (qL[2] * vL + pL) * ny,
                     (uL * (qL[3] + pL)) * nx + (vL * (qL[3] + pL)) * ny};
    double FnR[4] = {qR[1] * nx + qR[2] * ny, (qR[1] * uR + pR) * nx + (qR[1] * vR) * ny,
                     (qR[1] * vR) * nx + (qR[2] * vR + pR) * ny,
                     (uR * (qR[3] + pR)) * nx + (vR * (qR[3] + pR)) * ny};
    double S_L = unL - aL, tmp = unR - aR;
    if (tmp < S_L) S_L = tmp;
    double S_R = unL + aL;
    tmp = unR + aR;
    if (tmp > S_R) S_R = tmp;
    double num = pR - pL + rho_L * unL * (S_L - unL) - rho_R * unR * (S_R - unR);
    double den = rho_L * (S_L - unL) - rho_R * (S_R - unR);
    double S_star = num / den;
    if (S_L >= 0.0) {
        for (int c = 0; c < 4; ++c) f[c] = FnL[c];
    } else if (S_R <= 0.0) {
        for (int c = 0; c < 4; ++c) f[c] = FnR[c];
    } else if (S_star >= 0.0) {
        double coef = (S_L - unL) / (S_L - S_star);
        double rs = rho_L * coef;
        double us = uL + (S_star - unL) * nx, vs = vL + (S_star - unL) * ny;
        double Es = qL[3] / rho_L + (S_star - unL) * (S_star + pL / (rho_L * (S_L - unL)));
        double qs[4] = {rs, rs * us, rs * vs, rs * Es};
        for (int c = 0; c < 4; ++c) f[c] = FnL[c] + S_L * (qs[c] - qL[c]);
    } else {
        double coef = (S_R - unR) / (S_R - S_star);
        double rs = rho_R * coef;
        double us = uR + (S_star - unR) * nx, vs = vR + (S_star - unR) * ny;
        double Es = qR[3] / rho_R + (S_star - unR) * (S_star + pR / (rho_R * (S_R - unR)));
        double qs[4] = {rs, rs * us, rs * vs, rs * Es};
        for (int c = 0; c < 4; ++c) f[c] = FnR[c] + S_R * (qs[c] - qR[c]);
    }
}

static void sw_state(const double *prm, const double *q, double *Fx, double *Fy, double *speed) {
    /* src/local_mesh_2d_gpu_sw.mojo:232-262 */
    double g = prm[0], hmin = prm[1];
    double h = q[0] < hmin ? hmin : q[0];
    double u = q[1] / h, v = q[2] / h, p = 0.5 * g * h * h;
    Fx[0] = q[1]; Fx[1] = q[1] * u + p; Fx[2] = q[1] * v;
    Fy[0] = q[2]; Fy[1] = q[2] * u; Fy[2] = q[2] * v + p;
    if (speed) *speed = sqrt(u * u + v * v) + sqrt(g * h);
}

static void sw_hll(const double *prm, const double *qL, const double *qR, double nx, double ny,
                   double *f) {
    /* src/local_mesh_2d_gpu_sw.mojo:360-426 */
    double g = prm[0], hmin = prm[1];
    double hL = qL[0] < hmin ? hmin : qL[0], hR = qR[0] < hmin ? hmin : qR[0];
    double uL = qL[1] / hL, vL = qL[2] / hL, uR = qR[1] / hR, vR = qR[2] / hR;
    double pL = 0.5 * g * hL * hL, pR = 0.5 * g * hR * hR;
    double cL = sqrt(g * hL), cR = sqrt(g * hR);
    double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
    double S_L = unL - cL, tmp = unR - cR;
    if (tmp < S_L) S_L = tmp;
    double S_R = unL + cL;
    tmp = unR + cR;
    if (tmp > S_R) S_R = tmp;
    double FxL[3] = {qL[1], qL[1] * uL + pL, qL[1] * vL}, FyL[3] = {qL[2], qL[2] * uL, qL[2] * vL + pL};
    double FxR[3] = {qR[1], qR[1] * uR + pR, qR[1] * vR}, FyR[3] = {qR[2], qR[2] * uR, qR[2] * vR + pR};
    double FnL[3], FnR[3];
    for (int c = 0; c < 3; ++c) {
        FnL[c] = FxL[c] * nx + FyL[c] * ny;
        FnR[c] = FxR[c] * nx + FyR[c] * ny;
    }
    if (S_L >= 0.0) {
        for (int c = 0; c < 3; ++c) f[c] = FnL[c];
    } else if (S_R <= 0.0) {
        for (int c = 0; c < 3; ++c) f[c] = FnR[c];
    } else {
        double inv = 1.0 / (S_R - S_L);
        for (int c = 0; c < 3; ++c)
            f[c] = (S_R * FnL[c] - S_L * FnR[c] + S_L * S_R * (qR[c] - qL[c])) * inv;
    }
}

static void glm_state(const double *prm, const double *q, double *Fx, double *Fy, double *speed) {
    /* src/local_mesh_2d_gpu_mhd_glm.mojo:334-364 (face) and 87-123 (volume) */
    double g = prm[0], rmin = prm[1], pmin = prm[2], ch = prm[3];
    double ch2 = ch * ch;
    double rho = q[0] < rmin ? rmin : q[0];
    double u = q[1] / rho, v = q[2] / rho;
    double Bx = q[3], By = q[4], E = q[5], psi = q[6];
    double BB = Bx * Bx + By * By;
    double ke = 0.5 * (q[1] * q[1] + q[2] * q[2]) / rho;
    double p = (g - 1.0) * (E - ke - 0.5 * BB);
    if (p < pmin) p = pmin;
    double ps = p + 0.5 * BB;
    Fx[0] = q[1]; Fx[1] = q[1] * u + ps - Bx * Bx; Fx[2] = q[1] * v - Bx * By; Fx[3] = psi;
    Fx[4] = u * By - v * Bx; Fx[5] = (E + ps) * u - Bx * (u * Bx + v * By); Fx[6] = ch2 * Bx;
    Fy[0] = q[2]; Fy[1] = q[2] * u - By * Bx; Fy[2] = q[2] * v + ps - By * By; Fy[3] = v * Bx - u * By;
    Fy[4] = psi; Fy[5] = (E + ps) * v - By * (u * Bx + v * By); Fy[6] = ch2 * By;
    if (speed) {
        double cs2 = g * p / rho, ca2 = BB / rho, s = cs2 + ca2;
        double disc = s * s - 4.0 * cs2 * (Bx * Bx) / rho;
        if (disc < 0.0) disc = 0.0;
        double cf = sqrt(0.5 * (s + sqrt(disc)));
        double sp = sqrt(u * u + v * v) + cf;
        if (ch > sp) sp = ch;
        *speed = sp;
    }
}

static void maxwell_state(const double *prm, const double *q, double *Fx, double *Fy) {
    /* src/local_mesh_2d_gpu_maxwell.mojo:79-99 */
    double c2 = prm[0] * prm[0];
    Fx[0] = 0.0; Fx[1] = c2 * q[5]; Fx[2] = -c2 * q[4]; Fx[3] = 0.0; Fx[4] = -q[2]; Fx[5] = q[1];
    Fy[0] = -c2 * q[5]; Fy[1] = 0.0; Fy[2] = c2 * q[3]; Fy[3] = q[2]; Fy[4] = 0.0; Fy[5] = -q[0];
}

/* Nodal volume flux of the 2D module physics (x and y parts). */
int orc_module_volume_flux(int phys, const double *prm, const double *q, double *Fx, double *Fy);
int orc_module_volume_flux(int phys, const double *prm, const double *q, double *Fx, double *Fy) {
    switch (phys) {
    case ORC_PHYS_EULER2D: e2d_state(prm, q, Fx, Fy, NULL, NULL, NULL); return 0;
    case ORC_PHYS_SW2D: sw_state(prm, q, Fx, Fy, NULL); return 0;
    case ORC_PHYS_MHDGLM2D: glm_state(prm, q, Fx, Fy, NULL); return 0;
    case ORC_PHYS_MAXWELL2D: maxwell_state(prm, q, Fx, Fy); return 0;
    case ORC_PHYS_ADVECTION2D: Fx[0] = prm[0] * q[0]; Fy[0] = prm[1] * q[0]; return 0;
    }
    return -1;
}

/* Ghost state of the 2D module face kernels. */
static void module_ghost(int phys, const double *prm, const double *qL, int bc, double nx, double ny,
                         double *qR) {
    int nc = orc_phys_nc(phys);
    if (bc == 3) { /* BC_INFLOW */
        const double *in = NULL;
        if (phys == ORC_PHYS_EULER2D) in = prm + 6;
        if (phys == ORC_PHYS_SW2D) in = prm + 3;
        if (phys == ORC_PHYS_MHDGLM2D) in = prm + 5;
        if (phys == ORC_PHYS_MAXWELL2D) in = prm + 7;
        for (int c = 0; c < nc; ++c) qR[c] = in[c];
        return;
    }
    if (bc == 1) { /* BC_WALL */
        if (phys == ORC_PHYS_MAXWELL2D) {
            double En = qL[0] * nx + qL[1] * ny, Bn = qL[3] * nx + qL[4] * ny;
            qR[0] = 2.0 * En * nx - qL[0];
            qR[1] = 2.0 * En * ny - qL[1];
            qR[2] = -qL[2];
            qR[3] = qL[3] - 2.0 * Bn * nx;
            qR[4] = qL[4] - 2.0 * Bn * ny;
            qR[5] = qL[5];
            return;
        }
        for (int c = 0; c < nc; ++c) qR[c] = qL[c];
        double mn = qL[1] * nx + qL[2] * ny;
        qR[1] = qL[1] - 2.0 * mn * nx;
        qR[2] = qL[2] - 2.0 * mn * ny;
        if (phys == ORC_PHYS_MHDGLM2D) {
            double Bn = qL[3] * nx + qL[4] * ny;
            qR[3] = qL[3] - 2.0 * Bn * nx;
            qR[4] = qL[4] - 2.0 * Bn * ny;
            qR[6] = -qL[6];
        }
        return;
    }
    for (int c = 0; c < nc; ++c) qR[c] = qL[c]; /* BC_OUTFLOW / default */
}

int orc_face_flux_2d(int phys, const double *prm, const double *ql, const double *qr_in, int bc,
                     double nx, double ny, double *f) {
    if (phys == ORC_PHYS_ADVECTION2D) {
        /* src/local_mesh_2d_gpu_advection.mojo:270-300 */
        double vn = prm[0] * nx + prm[1] * ny;
        if (bc == 0)
            f[0] = vn >= 0.0 ? vn * ql[0] : vn * qr_in[0];
        else if (bc == 3)
            f[0] = vn >= 0.0 ? vn * ql[0] : vn * prm[2];
        else
            f[0] = vn >= 0.0 ? vn * ql[0] : 0.0;
        return 0;
    }
    int nc = orc_phys_nc(phys);
    if (nc < 0) return -1;
    double qr[8] = {0};
    if (bc == 0)
        for (int c = 0; c < nc; ++c) qr[c] = qr_in[c];
    else
        module_ghost(phys, prm, ql, bc, nx, ny, qr);
    if (phys == ORC_PHYS_EULER2D && prm[3] != 0.0) {
        e2d_hllc(prm, ql, qr, nx, ny, f);
        return 0;
    }
    if (phys == ORC_PHYS_SW2D && prm[2] != 0.0) {
        sw_hll(prm, ql, qr, nx, ny, f);
        return 0;
    }
    if (phys == ORC_PHYS_MAXWELL2D) {
        /* src/local_mesh_2d_gpu_maxwell.mojo:311-334 */
        double c2 = prm[0] * prm[0], al = prm[0];
        double FL[6] = {c2 * (-ql[5] * ny), c2 * (ql[5] * nx), c2 * (ql[3] * ny - ql[4] * nx),
                        ny * ql[2], -nx * ql[2], nx * ql[1] - ny * ql[0]};
        double FR[6] = {c2 * (-qr[5] * ny), c2 * (qr[5] * nx), c2 * (qr[3] * ny - qr[4] * nx),
                        ny * qr[2], -nx * qr[2], nx * qr[1] - ny * qr[0]};
        for (int c = 0; c < 6; ++c) f[c] = 0.5 * (FL[c] + FR[c]) - 0.5 * al * (qr[c] - ql[c]);
        return 0;
    }
    /* Rusanov with |u|+c bound (Euler2D, SW, GLM-MHD). */
    double FxL[8] = {0}, FyL[8] = {0}, FxR[8] = {0}, FyR[8] = {0}, sL = 0.0, sR = 0.0;
    if (phys == ORC_PHYS_EULER2D) {
        e2d_state(prm, ql, FxL, FyL, &sL, NULL, NULL);
        e2d_state(prm, qr, FxR, FyR, &sR, NULL, NULL);
    } else if (phys == ORC_PHYS_SW2D) {
        sw_state(prm, ql, FxL, FyL, &sL);
        sw_state(prm, qr, FxR, FyR, &sR);
    } else {
        glm_state(prm, ql, FxL, FyL, &sL);
        glm_state(prm, qr, FxR, FyR, &sR);
    }
    double al = sL > sR ? sL : sR;
    for (int c = 0; c < nc; ++c)
        f[c] = 0.5 * ((FxL[c] + FxR[c]) * nx + (FyL[c] + FyR[c]) * ny) - 0.5 * al * (qr[c] - ql[c]);
    return 0;
}
