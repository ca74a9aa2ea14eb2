// physics.hpp -- pointwise physics of the DG right-hand side as
// __host__ __device__ functors, one per physics kind.
//
// Interface (the Physics trait of src/solver.mojo:97-145):
//   NC                        number of conserved components
//   flux<ND>(q, F)            F[d*NC + c], world-frame flux, d < ND
//   numerical_flux(ql,qr,n,f) two-state flux, n points from ql's element
//   ghost(q,bc,n,g)           ghost state of a boundary face (n outward);
//                             boundary_flux = numerical_flux(q, ghost(q))
//   source(q, s)              pointwise source (kSource)
//   limit(q)                  post-stage positivity fix (kLimit)
// Trait physics (Advection, Euler) run in 3D and, with nz = 0, on
// triangles; the 2D module physics (Euler2D, SW, GLM-MHD, Maxwell,
// Advection2D) follow their reference kernels' inline physics.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#define MXM_HD __host__ __device__ __forceinline__

namespace mxm {

enum : int { BC_INTERIOR = 0, BC_WALL = 1, BC_OUTFLOW = 2, BC_INFLOW = 3 };

template <class R>
MXM_HD R rmax(R a, R b) { return a > b ? a : b; }
template <class R>
MXM_HD R rmin(R a, R b) { return a < b ? a : b; }
template <class R>
MXM_HD R rabs(R a) { return a >= R(0) ? a : -a; }
// square root in the precision of the argument
MXM_HD float msqrt(float x) { return sqrtf(x); }
// max that keeps a NaN in its first argument (the CFL reduction must not
// hide a broken state behind a constant speed)
template <class R>
MXM_HD R nmax(R a, R b) { return (a > b || a != a) ? a : b; }
// FP64: the hardware reciprocal square root refined by Goldschmidt's
// iteration (g -> sqrt(x), h -> 1/(2 sqrt(x))) and two Newton corrections
// of g, within an ulp or so of the rounded root.  The fluxes take roots of
// positive normal numbers or of an exact zero, so the operand scaling of
// the library square root is not needed; x = 0 (where the reciprocal root
// is infinite) returns 0.
__device__ __forceinline__ double msqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  g = fma(d, h, g);
  return x > 0.0 ? g : 0.0;
}

// ------------------------------------------------------------ Advection --
// src/advection.mojo:60-119
template <class R>
struct Advection {
  static constexpr int NC = 1;
  static constexpr bool kSource = false, kLimit = false;
  R v[3];
  R inflow;
  MXM_HD static Advection from(const double *p) { return {{R(p[0]), R(p[1]), R(p[2])}, R(p[3])}; }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    for (int d = 0; d < ND; ++d) F[d] = v[d] * q[0];
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    R a = rabs(vn);
    f[0] = R(0.5) * ((vn + a) * ql[0] + (vn - a) * qr[0]);
  }
  // WALL / OUTFLOW: zero Dirichlet on the inflowing half; INFLOW: user value
  // (src/advection.mojo:98-112)
  MXM_HD void ghost(const R *, int bc, const R *, R *g) const { g[0] = bc == BC_INFLOW ? inflow : R(0); }
  MXM_HD void source(const R *, R *s) const { s[0] = R(0); }
  // largest signal speed at a node, over all directions (CFL reduction)
  MXM_HD R max_speed(const R *) const { return msqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// ---------------------------------------------------------------- Euler --
// src/euler.mojo:44-792.  FLUX: 0 Rusanov, 1 Roe, 2 HLLE, 3 HLLEC;
// EFIX: Harten-Hyman entropy fix on the wave-based solvers.
template <class R, int FLUX, int EFIX = 0>
struct Euler {
  static constexpr int NC = 5;
  static constexpr bool kSource = true, kLimit = true;
  static constexpr bool efix = EFIX != 0;
  R gamma, rmin_, pmin_;
  R g[3];
  R inflow[5];
  double gamma_d, rmin_d, pmin_d;  // FP64 copies for the positivity rescue
  MXM_HD static Euler from(const double *p) {
    Euler e;
    e.gamma = R(p[0]);
    e.rmin_ = R(p[1]);
    e.pmin_ = R(p[2]);
    e.gamma_d = p[0], e.rmin_d = p[1], e.pmin_d = p[2];
    for (int d = 0; d < 3; ++d) e.g[d] = R(p[5 + d]);
    for (int c = 0; c < 5; ++c) e.inflow[c] = R(p[8 + c]);
    return e;
  }
  MXM_HD R rho(const R *q) const { return q[0] > rmin_ ? q[0] : rmin_; }
  MXM_HD R pres(const R *q) const {
    R r = rho(q);
    R ke = R(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / r;
    R p = (gamma - R(1)) * (q[4] - ke);
    return p > pmin_ ? p : pmin_;
  }
  // Floored primitives (src/euler.mojo:44-56) from one reciprocal of the
  // density: FP64 division is a long instruction sequence on CDNA, so the
  // reference's repeated m/rho become multiplications by 1/rho.
  struct Prim {
    R r, ir, u[3], p;
  };
  MXM_HD Prim prim(const R *q) const {
    Prim s;
    s.r = q[0] > rmin_ ? q[0] : rmin_;
    s.ir = R(1) / s.r;
    for (int d = 0; d < 3; ++d) s.u[d] = q[1 + d] * s.ir;
    R ke = R(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * s.ir;
    R p = (gamma - R(1)) * (q[4] - ke);
    s.p = p > pmin_ ? p : pmin_;
    return s;
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    const Prim s = prim(q);
    for (int d = 0; d < ND; ++d) {
      R *Fd = F + d * 5;
      Fd[0] = q[1 + d];
      Fd[1] = q[1 + d] * s.u[0];
      Fd[2] = q[1 + d] * s.u[1];
      Fd[3] = q[1 + d] * s.u[2];
      Fd[1 + d] += s.p;
      Fd[4] = s.u[d] * (q[4] + s.p);
    }
  }
  // x-direction flux of a face-aligned state
  MXM_HD static void flux1(const R *q, const Prim &s, R *f) {
    f[0] = q[1];
    f[1] = q[1] * s.u[0] + s.p;
    f[2] = q[2] * s.u[0];
    f[3] = q[3] * s.u[0];
    f[4] = s.u[0] * (q[4] + s.p);
  }
  struct Frame {
    R n[3], t[3], b[3];
  };
  // tangent / binormal, src/euler.mojo:94-118
  MXM_HD static Frame frame(const R *n) {
    Frame F;
    // scalars, not an array: a select on the perturbed component would
    // otherwise become a dynamically indexed private array (scratch)
    const bool big = rabs(n[0]) >= R(0.9);
    R ax = big ? n[0] : n[0] + R(1);
    R ay = big ? n[1] + R(1) : n[1];
    R az = n[2];
    const R pr = n[0] * ax + n[1] * ay + n[2] * az;
    ax -= n[0] * pr;
    ay -= n[1] * pr;
    az -= n[2] * pr;
    const R imag = R(1) / msqrt(ax * ax + ay * ay + az * az);
    for (int d = 0; d < 3; ++d) F.n[d] = n[d];
    F.t[0] = ax * imag;
    F.t[1] = ay * imag;
    F.t[2] = az * imag;
    F.b[0] = n[1] * F.t[2] - n[2] * F.t[1];
    F.b[1] = n[2] * F.t[0] - n[0] * F.t[2];
    F.b[2] = n[0] * F.t[1] - n[1] * F.t[0];
    return F;
  }
  MXM_HD static void to_local(const Frame &F, const R *q, R *r) {
    r[0] = q[0];
    r[4] = q[4];
    r[1] = F.n[0] * q[1] + F.n[1] * q[2] + F.n[2] * q[3];
    r[2] = F.t[0] * q[1] + F.t[1] * q[2] + F.t[2] * q[3];
    r[3] = F.b[0] * q[1] + F.b[1] * q[2] + F.b[2] * q[3];
  }
  MXM_HD static void to_world(const Frame &F, const R *r, R *q) {
    q[0] = r[0];
    q[4] = r[4];
    for (int d = 0; d < 3; ++d) q[1 + d] = F.n[d] * r[1] + F.t[d] * r[2] + F.b[d] * r[3];
  }
  // Roe averages, src/euler.mojo:161-176 (returns u, v, w, H, a):
  // m/msqrt(rho) = msqrt(rho) u, so every term shares one reciprocal.
  MXM_HD void roe_avg(const R *ql, const R *qr, const Prim &L, const Prim &Rt, R *o) const {
    const R sl = msqrt(L.r), sr = msqrt(Rt.r), is = R(1) / (sl + sr);
    for (int d = 0; d < 3; ++d) o[d] = (sl * L.u[d] + sr * Rt.u[d]) * is;
    o[3] = (sl * ((ql[4] + L.p) * L.ir) + sr * ((qr[4] + Rt.p) * Rt.ir)) * is;
    R u2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    o[4] = msqrt(rmax((gamma - R(1)) * (o[3] - R(0.5) * u2), R(0)));
  }
  // Harten-Hyman entropy fix, src/euler.mojo:207-288.  wave[m*3 + w].
  MXM_HD void entropy_fixed(const R *ql, const R *qr, const R *s, const R *wave, R *amdq,
                            R *apdq) const {
    const R gm = gamma;
    R rl = rho(ql), pl = pres(ql);
    R s0 = ql[1] / rl - msqrt(gm * pl / rl);
    for (int m = 0; m < 5; ++m) amdq[m] = R(0);
    if (!(s0 > R(0) && s[0] > R(0))) {
      R a[5];
      for (int m = 0; m < 5; ++m) a[m] = ql[m] + wave[m * 3];
      a[0] = rmax(a[0], rmin_);
      R ke = R(0.5) * (a[1] * a[1] + a[2] * a[2] + a[3] * a[3]) / a[0];
      R p1 = rmax((gm - R(1)) * (a[4] - ke), pmin_);
      R s1 = a[1] / a[0] - msqrt(gm * p1 / a[0]);
      R sf = (s0 < R(0) && s1 > R(0)) ? s0 * (s1 - s[0]) / (s1 - s0) : (s[0] < R(0) ? s[0] : R(0));
      for (int m = 0; m < 5; ++m) amdq[m] = sf * wave[m * 3];
      if (s[1] < R(0)) {
        for (int m = 0; m < 5; ++m) amdq[m] += s[1] * wave[m * 3 + 1];
        R rr = rho(qr), prr = pres(qr);
        R s3 = qr[1] / rr + msqrt(gm * prr / rr);
        R b[5];
        for (int m = 0; m < 5; ++m) b[m] = qr[m] - wave[m * 3 + 2];
        b[0] = rmax(b[0], rmin_);
        R ke2 = R(0.5) * (b[1] * b[1] + b[2] * b[2] + b[3] * b[3]) / b[0];
        R p2 = rmax((gm - R(1)) * (b[4] - ke2), pmin_);
        R s2 = b[1] / b[0] + msqrt(gm * p2 / b[0]);
        R sf2 = (s2 < R(0) && s3 > R(0)) ? s2 * (s3 - s[2]) / (s3 - s2) : (s[2] < R(0) ? s[2] : R(0));
        for (int m = 0; m < 5; ++m) amdq[m] += sf2 * wave[m * 3 + 2];
      }
    }
    for (int m = 0; m < 5; ++m) {
      R t = R(0);
      for (int w = 0; w < 3; ++w) t += s[w] * wave[m * 3 + w];
      apdq[m] = t - amdq[m];
    }
  }
  // face-aligned flux (rl, rr already rotated); src/euler.mojo:612-683
  MXM_HD void local_flux(const R *rl, const R *rr, R *rf) const {
    const Prim L = prim(rl), Rt = prim(rr);
    const R dl = L.r, dr = Rt.r, pl = L.p, pr = Rt.p, ul = L.u[0], ur = Rt.u[0];
    if constexpr (FLUX == 0) {
      R cl = msqrt(gamma * pl * L.ir), cr = msqrt(gamma * pr * Rt.ir);
      R cm = rmax(rabs(ul) + cl, rabs(ur) + cr);
      rf[0] = R(0.5) * (rl[1] + rr[1] + cm * (dl - dr));
      rf[1] = R(0.5) * (rl[1] * ul + pl + rr[1] * ur + pr + cm * (rl[1] - rr[1]));
      rf[2] = R(0.5) * (rl[2] * ul + rr[2] * ur + cm * (rl[2] - rr[2]));
      rf[3] = R(0.5) * (rl[3] * ul + rr[3] * ur + cm * (rl[3] - rr[3]));
      rf[4] = R(0.5) * ((rl[4] + pl) * ul + (rr[4] + pr) * ur + cm * (rl[4] - rr[4]));
    } else if constexpr (FLUX == 3 && !efix) {
      // HLLEC without the entropy fix in its upwind-side form:
      // F = F_K + s_K^- (U*_K - U_K) on the side K of the contact (s_L^- =
      // min(s_L, 0) on the left, s_R^+ = max(s_R, 0) on the right).  The
      // star states satisfy the jump conditions across each of the three
      // waves, so this equals the wave sum of src/euler.mojo:415-520
      // (1/2 (F_L + F_R) - 1/2 sum_w |s_w| W_w) up to rounding, with one
      // star state and one physical flux instead of two.  In FP32 its
      // rounding is larger than the wave sum's: the 128^2 P=3 vortex at
      // T=2 is 5.2e-6 from FP64 with it against 3.0e-6 with the wave sum
      // (stochastically rounded state), for 3 % less time per stage.
      R ra[5];
      roe_avg(rl, rr, L, Rt, ra);
      const R s0 = rmin(ul - msqrt(gamma * pl * L.ir), ra[0] - ra[4]);
      const R s2 = rmax(ur + msqrt(gamma * pr * Rt.ir), ra[0] + ra[4]);
      const R Al = dl * (s0 - ul), Ar = dr * (s2 - ur);
      const R s1 = (pr - pl + Al * ul - Ar * ur) / (Al - Ar);
      const bool left = s1 >= R(0);
      const R *q = left ? rl : rr;
      const Prim &K = left ? L : Rt;
      const R sK = left ? s0 : s2, AK = left ? Al : Ar;
      const R wK = left ? rmin(s0, R(0)) : rmax(s2, R(0));
      const R lm = AK / (sK - s1);
      const R h[5] = {lm, lm * s1, lm * K.u[1], lm * K.u[2],
                      lm * (q[4] * K.ir + (s1 - K.u[0]) * (s1 + K.p / AK))};
      R fk[5];
      flux1(q, K, fk);
      for (int m = 0; m < 5; ++m) rf[m] = fk[m] + wK * (h[m] - q[m]);
    } else {
      R s[3], wave[15];
      [[maybe_unused]] R amdq[5], apdq[5];
      R ra[5];
      roe_avg(rl, rr, L, Rt, ra);
      if constexpr (FLUX == 1) {
        // Roe, src/euler.mojo:296-361
        const R g1 = gamma - R(1);
        R d[5];
        for (int m = 0; m < 5; ++m) d[m] = rr[m] - rl[m];
        R u = ra[0], v = ra[1], w = ra[2], H = ra[3], a = ra[4];
        R u2 = u * u + v * v + w * w;
        const R ia = R(1) / a;
        R a4 = g1 * (ia * ia) * ((H - u2) * d[0] + u * d[1] + v * d[2] + w * d[3] - d[4]);
        R a2 = d[2] - v * d[0], a3 = d[3] - w * d[0];
        R a5 = (d[1] + (a - u) * d[0] - a * a4) * (R(0.5) * ia);
        R a1 = d[0] - a4 - a5;
        wave[0] = a1, wave[3] = a1 * (u - a), wave[6] = a1 * v, wave[9] = a1 * w;
        wave[12] = a1 * (H - u * a);
        wave[1] = a4, wave[4] = a4 * u, wave[7] = a4 * v + a2, wave[10] = a4 * w + a3;
        wave[13] = a4 * R(0.5) * u2 + a2 * v + a3 * w;
        wave[2] = a5, wave[5] = a5 * (u + a), wave[8] = a5 * v, wave[11] = a5 * w;
        wave[14] = a5 * (H + u * a);
        s[0] = u - a, s[1] = u, s[2] = u + a;
      } else {
        s[0] = rmin(ul - msqrt(gamma * pl * L.ir), ra[0] - ra[4]);
        s[2] = rmax(ur + msqrt(gamma * pr * Rt.ir), ra[0] + ra[4]);
        if constexpr (FLUX == 2) {
          // HLLE, src/euler.mojo:364-412
          s[1] = R(0);
          R fl[5], fr[5];
          flux1(rl, L, fl);
          flux1(rr, Rt, fr);
          const R iden = R(1) / (s[0] - s[2]);
          for (int m = 0; m < 5; ++m) {
            R qh = (fr[m] - fl[m] - s[2] * rr[m] + s[0] * rl[m]) * iden;
            wave[m * 3] = qh - rl[m];
            wave[m * 3 + 1] = R(0);
            wave[m * 3 + 2] = rr[m] - qh;
          }
        } else {
          // HLLEC, src/euler.mojo:415-485
          const R Al = dl * (s[0] - ul), Ar = dr * (s[2] - ur);
          s[1] = (pr - pl + Al * ul - Ar * ur) / (Al - Ar);
          // two reciprocals for the four quotients Al/(s0-s1), Ar/(s2-s1),
          // pl/Al, pr/Ar
          const R d0 = s[0] - s[1], d2 = s[2] - s[1];
          const R i02 = R(1) / (d0 * d2), iA = R(1) / (Al * Ar);
          const R lm = Al * d2 * i02, rm = Ar * d0 * i02;
          R hl[5] = {lm, lm * s[1], lm * L.u[1], lm * L.u[2],
                     lm * (rl[4] * L.ir + (s[1] - ul) * (s[1] + pl * Ar * iA))};
          R hr[5] = {rm, rm * s[1], rm * Rt.u[1], rm * Rt.u[2],
                     rm * (rr[4] * Rt.ir + (s[1] - ur) * (s[1] + pr * Al * iA))};
          for (int m = 0; m < 5; ++m) {
            wave[m * 3] = hl[m] - rl[m];
            wave[m * 3 + 1] = hr[m] - hl[m];
            wave[m * 3 + 2] = rr[m] - hr[m];
          }
        }
      }
      R fl[5], fr[5];
      flux1(rl, L, fl);
      flux1(rr, Rt, fr);
      if constexpr (efix) {
        entropy_fixed(rl, rr, s, wave, amdq, apdq);
        for (int m = 0; m < 5; ++m) rf[m] = R(0.5) * (fl[m] + fr[m]) - R(0.5) * (apdq[m] - amdq[m]);
      } else {
        // apdq - amdq = sum_w |s_w| wave_w (src/euler.mojo:184-199, 495-520)
        const R a0 = rabs(s[0]), a1 = rabs(s[1]), a2 = rabs(s[2]);
        for (int m = 0; m < 5; ++m)
          rf[m] = R(0.5) * (fl[m] + fr[m]) - R(0.5) * (a0 * wave[m * 3] + a1 * wave[m * 3 + 1] + a2 * wave[m * 3 + 2]);
      }
    }
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    if constexpr (FLUX == 0) {
      // Rusanov in the world frame: the rotated form below gives the same
      // flux (normal velocity and pressure are frame invariant, the
      // momentum terms rotate back), up to rounding, without the frame
      const Prim L = prim(ql), Rt = prim(qr);
      const R unl = L.u[0] * n[0] + L.u[1] * n[1] + L.u[2] * n[2];
      const R unr = Rt.u[0] * n[0] + Rt.u[1] * n[1] + Rt.u[2] * n[2];
      const R cl = msqrt(gamma * L.p * L.ir), cr = msqrt(gamma * Rt.p * Rt.ir);
      const R cm = rmax(rabs(unl) + cl, rabs(unr) + cr);
      const R mnl = ql[1] * n[0] + ql[2] * n[1] + ql[3] * n[2], mnr = qr[1] * n[0] + qr[2] * n[1] + qr[3] * n[2];
      f[0] = R(0.5) * (mnl + mnr + cm * (L.r - Rt.r));
      for (int d = 0; d < 3; ++d)
        f[1 + d] = R(0.5) * (ql[1 + d] * unl + L.p * n[d] + qr[1 + d] * unr + Rt.p * n[d] + cm * (ql[1 + d] - qr[1 + d]));
      f[4] = R(0.5) * ((ql[4] + L.p) * unl + (qr[4] + Rt.p) * unr + cm * (ql[4] - qr[4]));
      return;
    }
    R rl[5], rr[5], rf[5];
    if constexpr (ND == 2) {
      // n = (nx, ny, 0): the frame (n, (-ny, nx, 0), (0, 0, 1)).  The flux
      // is invariant under the choice of tangent basis (only |v_t| and the
      // tangential components as vectors enter), so this equals the
      // reference's constructed frame up to rounding.
      const R nx = n[0], ny = n[1];
      rl[0] = ql[0], rl[4] = ql[4], rr[0] = qr[0], rr[4] = qr[4];
      rl[1] = nx * ql[1] + ny * ql[2], rl[2] = nx * ql[2] - ny * ql[1], rl[3] = ql[3];
      rr[1] = nx * qr[1] + ny * qr[2], rr[2] = nx * qr[2] - ny * qr[1], rr[3] = qr[3];
      local_flux(rl, rr, rf);
      f[0] = rf[0], f[4] = rf[4], f[3] = rf[3];
      f[1] = nx * rf[1] - ny * rf[2];
      f[2] = ny * rf[1] + nx * rf[2];
    } else {
      Frame F = frame(n);
      to_local(F, ql, rl);
      to_local(F, qr, rr);
      local_flux(rl, rr, rf);
      to_world(F, rf, f);
    }
  }
  // ghost states, src/euler.mojo:705-730: WALL reflects the normal
  // momentum, INFLOW is the stored state, OUTFLOW copies
  // (per-component selects: branches here become scratch-indexed stores)
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    const bool wall = bc == BC_WALL, in = bc == BC_INFLOW;
    const R mn = q[1] * n[0] + q[2] * n[1] + q[3] * n[2];
    g[0] = in ? inflow[0] : q[0];
    g[4] = in ? inflow[4] : q[4];
#pragma unroll
    for (int d = 0; d < 3; ++d) g[1 + d] = in ? inflow[1 + d] : (wall ? q[1 + d] - R(2) * mn * n[d] : q[1 + d]);
  }
  // gravity, src/euler.mojo:738-747
  MXM_HD bool source_active() const { return g[0] != R(0) || g[1] != R(0) || g[2] != R(0); }
  MXM_HD void source(const R *q, R *s) const {
    s[0] = R(0);
    s[1] = q[0] * g[0];
    s[2] = q[0] * g[1];
    s[3] = q[0] * g[2];
    s[4] = q[1] * g[0] + q[2] * g[1] + q[3] * g[2];
  }
  // |u| + c (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    const Prim s = prim(q);
    return msqrt(s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]) + msqrt(gamma * s.p * s.ir);
  }
  // NaN-safe positivity rescue, src/euler.mojo:758-792, in FP64 whatever
  // the state precision (the stage kernels pass an FP64 copy of the
  // combined node).  Both divisions are on the paths that rescue only: the
  // pressure test p > pmin, p = (gamma - 1) (E - m^2 / (2 rho)), is taken
  // multiplied through by rho > 0 (a NaN or infinite energy or momentum
  // fails it as before).
  MXM_HD void limit(double *q) const {
    if (!(q[0] > rmin_d)) {
      q[0] = rmin_d;
      q[1] = q[2] = q[3] = 0.0;
      q[4] = pmin_d / (gamma_d - 1.0);
      return;
    }
    double m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (!(m2 >= 0.0)) {
      q[1] = q[2] = q[3] = 0.0;
      q[4] = pmin_d / (gamma_d - 1.0);
      return;
    }
    if (!((gamma_d - 1.0) * (q[4] * q[0] - 0.5 * m2) > pmin_d * q[0])) q[4] = 0.5 * m2 / q[0] + pmin_d / (gamma_d - 1.0);
  }
};

// ---------------------------------------------------------------------------
// 2D module physics.  Boundary faces build the module's ghost state and call
// the same face flux; flux<2> is the nodal flux of the module's
// vol_lift_combine_rk kernel.
// ---------------------------------------------------------------------------

// src/local_mesh_2d_gpu_euler.mojo; HLLC when HLLC == 1
template <class R, int HLLC>
struct Euler2D {
  static constexpr int NC = 4;
  static constexpr bool kSource = true, kLimit = false;
  R gamma, rmin_, pmin_, gx, gy;
  R inflow[4];
  MXM_HD static Euler2D from(const double *p) {
    Euler2D e;
    e.gamma = R(p[0]), e.rmin_ = R(p[1]), e.pmin_ = R(p[2]), e.gx = R(p[4]), e.gy = R(p[5]);
    for (int c = 0; c < 4; ++c) e.inflow[c] = R(p[6 + c]);
    return e;
  }
  // primitive state with floors (volume and face kernels use the same)
  MXM_HD void prim(const R *q, R &r, R &u, R &v, R &p) const {
    r = q[0] < rmin_ ? rmin_ : q[0];
    const R ir = R(1) / r;
    u = q[1] * ir;
    v = q[2] * ir;
    R ke = R(0.5) * r * (u * u + v * v);
    p = (gamma - R(1)) * (q[3] - ke);
    if (p < pmin_) p = pmin_;
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    R r, u, v, p;
    prim(q, r, u, v, p);
    F[0] = q[1], F[1] = q[1] * u + p, F[2] = q[1] * v, F[3] = u * (q[3] + p);
    F[4] = q[2], F[5] = q[2] * u, F[6] = q[2] * v + p, F[7] = v * (q[3] + p);
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    const R nx = n[0], ny = n[1];
    R rl, ul, vl, pl, rr, ur, vr, pr;
    prim(ql, rl, ul, vl, pl);
    prim(qr, rr, ur, vr, pr);
    if constexpr (HLLC == 0) {
      R FL[8], FR[8];
      flux<2>(ql, FL);
      flux<2>(qr, FR);
      R sl = msqrt(ul * ul + vl * vl) + msqrt(gamma * pl * (R(1) / rl));
      R sr = msqrt(ur * ur + vr * vr) + msqrt(gamma * pr * (R(1) / rr));
      R al = sl > sr ? sl : sr;
      for (int c = 0; c < 4; ++c)
        f[c] = R(0.5) * ((FL[c] + FR[c]) * nx + (FL[4 + c] + FR[4 + c]) * ny) - R(0.5) * al * (qr[c] - ql[c]);
    } else {
      // src/local_mesh_2d_gpu_euler.mojo:466-545
      R al = msqrt(gamma * pl * (R(1) / rl)), ar = msqrt(gamma * pr * (R(1) / rr));
      R unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
      R FnL[4] = {ql[1] * nx + ql[2] * ny, (ql[1] * ul + pl) * nx + (ql[1] * vl) * ny,
                  (ql[1] * vl) * nx + (ql[2] * vl + pl) * ny, (ul * (ql[3] + pl)) * nx + (vl * (ql[3] + pl)) * ny};
      R FnR[4] = {qr[1] * nx + qr[2] * ny, (qr[1] * ur + pr) * nx + (qr[1] * vr) * ny,
                  (qr[1] * vr) * nx + (qr[2] * vr + pr) * ny, (ur * (qr[3] + pr)) * nx + (vr * (qr[3] + pr)) * ny};
      R SL = rmin(unl - al, unr - ar), SR = rmax(unl + al, unr + ar);
      R ss = (pr - pl + rl * unl * (SL - unl) - rr * unr * (SR - unr)) / (rl * (SL - unl) - rr * (SR - unr));
      if (SL >= R(0)) {
        for (int c = 0; c < 4; ++c) f[c] = FnL[c];
      } else if (SR <= R(0)) {
        for (int c = 0; c < 4; ++c) f[c] = FnR[c];
      } else {
        const bool left = ss >= R(0);
        const R *q = left ? ql : qr;
        const R *Fn = left ? FnL : FnR;
        R S = left ? SL : SR, d = left ? rl : rr, u = left ? ul : ur, v = left ? vl : vr;
        R un = left ? unl : unr, p = left ? pl : pr;
        R coef = (S - un) / (S - ss);
        R rs = d * coef;
        R us = u + (ss - un) * nx, vs = v + (ss - un) * ny;
        R Es = q[3] / d + (ss - un) * (ss + p / (d * (S - un)));
        R qs[4] = {rs, rs * us, rs * vs, rs * Es};
        for (int c = 0; c < 4; ++c) f[c] = Fn[c] + S * (qs[c] - q[c]);
      }
    }
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    const bool wall = bc == BC_WALL, in = bc == BC_INFLOW;
    const R mn = q[1] * n[0] + q[2] * n[1];
    g[0] = in ? inflow[0] : q[0];
    g[1] = in ? inflow[1] : (wall ? q[1] - R(2) * mn * n[0] : q[1]);
    g[2] = in ? inflow[2] : (wall ? q[2] - R(2) * mn * n[1] : q[2]);
    g[3] = in ? inflow[3] : q[3];
  }
  // gravity with the floored density, src/local_mesh_2d_gpu_euler.mojo:126-140
  MXM_HD bool source_active() const { return gx != R(0) || gy != R(0); }
  MXM_HD void source(const R *q, R *s) const {
    R r = q[0] < rmin_ ? rmin_ : q[0];
    s[0] = R(0), s[1] = r * gx, s[2] = r * gy, s[3] = q[1] * gx + q[2] * gy;
  }
  // |u| + c (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    R r, u, v, p;
    prim(q, r, u, v, p);
    return msqrt(u * u + v * v) + msqrt(gamma * p / r);
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/local_mesh_2d_gpu_sw.mojo; HLL when HLL == 1
template <class R, int HLL>
struct SW2D {
  static constexpr int NC = 3;
  static constexpr bool kSource = false, kLimit = false;
  R grav, hmin;
  R inflow[3];
  MXM_HD static SW2D from(const double *p) { return {R(p[0]), R(p[1]), {R(p[3]), R(p[4]), R(p[5])}}; }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    R h = q[0] < hmin ? hmin : q[0], ih = R(1) / h;
    R u = q[1] * ih, v = q[2] * ih, p = R(0.5) * grav * h * h;
    F[0] = q[1], F[1] = q[1] * u + p, F[2] = q[1] * v;
    F[3] = q[2], F[4] = q[2] * u, F[5] = q[2] * v + p;
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    const R nx = n[0], ny = n[1];
    R FL[6], FR[6];
    flux<2>(ql, FL);
    flux<2>(qr, FR);
    R hl = ql[0] < hmin ? hmin : ql[0], hr = qr[0] < hmin ? hmin : qr[0];
    const R ihl = R(1) / hl, ihr = R(1) / hr;  // the reciprocals flux<2> took
    R ul = ql[1] * ihl, vl = ql[2] * ihl, ur = qr[1] * ihr, vr = qr[2] * ihr;
    if constexpr (HLL == 0) {
      R sl = msqrt(ul * ul + vl * vl) + msqrt(grav * hl), sr = msqrt(ur * ur + vr * vr) + msqrt(grav * hr);
      R al = sl > sr ? sl : sr;
      for (int c = 0; c < 3; ++c)
        f[c] = R(0.5) * ((FL[c] + FR[c]) * nx + (FL[3 + c] + FR[3 + c]) * ny) - R(0.5) * al * (qr[c] - ql[c]);
    } else {
      R cl = msqrt(grav * hl), cr = msqrt(grav * hr);
      R unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
      R SL = rmin(unl - cl, unr - cr), SR = rmax(unl + cl, unr + cr);
      R FnL[3], FnR[3];
      for (int c = 0; c < 3; ++c) {
        FnL[c] = FL[c] * nx + FL[3 + c] * ny;
        FnR[c] = FR[c] * nx + FR[3 + c] * ny;
      }
      if (SL >= R(0)) {
        for (int c = 0; c < 3; ++c) f[c] = FnL[c];
      } else if (SR <= R(0)) {
        for (int c = 0; c < 3; ++c) f[c] = FnR[c];
      } else {
        R inv = R(1) / (SR - SL);
        for (int c = 0; c < 3; ++c) f[c] = (SR * FnL[c] - SL * FnR[c] + SL * SR * (qr[c] - ql[c])) * inv;
      }
    }
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    const bool wall = bc == BC_WALL, in = bc == BC_INFLOW;
    const R mn = q[1] * n[0] + q[2] * n[1];
    g[0] = in ? inflow[0] : q[0];
    g[1] = in ? inflow[1] : (wall ? q[1] - R(2) * mn * n[0] : q[1]);
    g[2] = in ? inflow[2] : (wall ? q[2] - R(2) * mn * n[1] : q[2]);
  }
  MXM_HD void source(const R *, R *s) const { s[0] = s[1] = s[2] = R(0); }
  // |u| + sqrt(g h) (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    const R h = q[0] < hmin ? hmin : q[0];
    return msqrt(q[1] * q[1] + q[2] * q[2]) / h + msqrt(grav * h);
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/local_mesh_2d_gpu_mhd_glm.mojo (GLM-MHD, Rusanov)
template <class R>
struct MHDGLM2D {
  static constexpr int NC = 7;
  static constexpr bool kSource = false, kLimit = false;
  R gamma, rmin_, pmin_, ch;
  R inflow[7];
  MXM_HD static MHDGLM2D from(const double *p) {
    MHDGLM2D e;
    e.gamma = R(p[0]), e.rmin_ = R(p[1]), e.pmin_ = R(p[2]), e.ch = R(p[3]);
    for (int c = 0; c < 7; ++c) e.inflow[c] = R(p[5 + c]);
    return e;
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    R sp;
    state(q, F, sp);
  }
  MXM_HD void state(const R *q, R *F, R &speed) const {
    R r = q[0] < rmin_ ? rmin_ : q[0], ir = R(1) / r;
    R u = q[1] * ir, v = q[2] * ir, Bx = q[3], By = q[4], E = q[5], psi = q[6];
    R BB = Bx * Bx + By * By;
    R ke = R(0.5) * (q[1] * q[1] + q[2] * q[2]) * ir;
    R p = (gamma - R(1)) * (E - ke - R(0.5) * BB);
    if (p < pmin_) p = pmin_;
    R ps = p + R(0.5) * BB, ch2 = ch * ch;
    F[0] = q[1], F[1] = q[1] * u + ps - Bx * Bx, F[2] = q[1] * v - Bx * By, F[3] = psi;
    F[4] = u * By - v * Bx, F[5] = (E + ps) * u - Bx * (u * Bx + v * By), F[6] = ch2 * Bx;
    F[7] = q[2], F[8] = q[2] * u - By * Bx, F[9] = q[2] * v + ps - By * By, F[10] = v * Bx - u * By;
    F[11] = psi, F[12] = (E + ps) * v - By * (u * Bx + v * By), F[13] = ch2 * By;
    R cs2 = gamma * p * ir, ca2 = BB * ir, s = cs2 + ca2;
    R disc = s * s - R(4) * cs2 * (Bx * Bx) * ir;
    if (disc < R(0)) disc = R(0);
    speed = rmax(msqrt(u * u + v * v) + msqrt(R(0.5) * (s + msqrt(disc))), ch);
  }
  // F.n of one side (the x and y fluxes of state() dotted with n) and its
  // Rusanov speed; 7 values live instead of 14
  MXM_HD void side(const R *q, R nx, R ny, R *Fn, R &speed) const {
    R r = q[0] < rmin_ ? rmin_ : q[0], ir = R(1) / r;
    R u = q[1] * ir, v = q[2] * ir, Bx = q[3], By = q[4], E = q[5], psi = q[6];
    R BB = Bx * Bx + By * By;
    R ke = R(0.5) * (q[1] * q[1] + q[2] * q[2]) * ir;
    R p = (gamma - R(1)) * (E - ke - R(0.5) * BB);
    if (p < pmin_) p = pmin_;
    R ps = p + R(0.5) * BB, ub = u * Bx + v * By;
    Fn[0] = q[1] * nx + q[2] * ny;
    Fn[1] = (q[1] * u + ps - Bx * Bx) * nx + (q[2] * u - By * Bx) * ny;
    Fn[2] = (q[1] * v - Bx * By) * nx + (q[2] * v + ps - By * By) * ny;
    Fn[3] = psi * nx + (v * Bx - u * By) * ny;
    Fn[4] = (u * By - v * Bx) * nx + psi * ny;
    Fn[5] = ((E + ps) * u - Bx * ub) * nx + ((E + ps) * v - By * ub) * ny;
    Fn[6] = ch * ch * (Bx * nx + By * ny);
    R cs2 = gamma * p * ir, ca2 = BB * ir, s = cs2 + ca2;
    R disc = s * s - R(4) * cs2 * (Bx * Bx) * ir;
    if (disc < R(0)) disc = R(0);
    speed = rmax(msqrt(u * u + v * v) + msqrt(R(0.5) * (s + msqrt(disc))), ch);
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R FL[7], FR[7], sl, sr;
    side(ql, n[0], n[1], FL, sl);
    side(qr, n[0], n[1], FR, sr);
    R al = sl > sr ? sl : sr;
    for (int c = 0; c < 7; ++c) f[c] = R(0.5) * (FL[c] + FR[c]) - R(0.5) * al * (qr[c] - ql[c]);
  }
  // WALL reflects m.n and B.n and flips psi (src/local_mesh_2d_gpu_mhd_glm.mojo:302-312)
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    const bool wall = bc == BC_WALL, in = bc == BC_INFLOW;
    const R mn = q[1] * n[0] + q[2] * n[1], Bn = q[3] * n[0] + q[4] * n[1];
    g[0] = in ? inflow[0] : q[0];
    g[1] = in ? inflow[1] : (wall ? q[1] - R(2) * mn * n[0] : q[1]);
    g[2] = in ? inflow[2] : (wall ? q[2] - R(2) * mn * n[1] : q[2]);
    g[3] = in ? inflow[3] : (wall ? q[3] - R(2) * Bn * n[0] : q[3]);
    g[4] = in ? inflow[4] : (wall ? q[4] - R(2) * Bn * n[1] : q[4]);
    g[5] = in ? inflow[5] : q[5];
    g[6] = in ? inflow[6] : (wall ? -q[6] : q[6]);
  }
  MXM_HD void source(const R *, R *s) const {
    for (int c = 0; c < 7; ++c) s[c] = R(0);
  }
  // max(|u| + the fast speed across B, c_h) (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    const R r = q[0] < rmin_ ? rmin_ : q[0];
    const R BB = q[3] * q[3] + q[4] * q[4];
    R p = (gamma - R(1)) * (q[5] - R(0.5) * (q[1] * q[1] + q[2] * q[2]) / r - R(0.5) * BB);
    if (p < pmin_) p = pmin_;
    return nmax(msqrt(q[1] * q[1] + q[2] * q[2]) / r + msqrt((gamma * p + BB) / r), ch);
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/local_mesh_2d_gpu_maxwell.mojo (Rusanov with alpha = c, PEC walls)
template <class R>
struct Maxwell2D {
  static constexpr int NC = 6;
  static constexpr bool kSource = true, kLimit = false;
  R c, J[3], M[3], inflow[6];
  MXM_HD static Maxwell2D from(const double *p) {
    Maxwell2D m;
    m.c = R(p[0]);
    for (int d = 0; d < 3; ++d) m.J[d] = R(p[1 + d]), m.M[d] = R(p[4 + d]);
    for (int k = 0; k < 6; ++k) m.inflow[k] = R(p[7 + k]);
    return m;
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    R c2 = c * c;
    F[0] = R(0), F[1] = c2 * q[5], F[2] = -c2 * q[4], F[3] = R(0), F[4] = -q[2], F[5] = q[1];
    F[6] = -c2 * q[5], F[7] = R(0), F[8] = c2 * q[3], F[9] = q[2], F[10] = R(0), F[11] = -q[0];
  }
  MXM_HD static void normal_flux(const R *q, R c2, R nx, R ny, R *F) {
    F[0] = c2 * (-q[5] * ny), F[1] = c2 * (q[5] * nx), F[2] = c2 * (q[3] * ny - q[4] * nx);
    F[3] = ny * q[2], F[4] = -nx * q[2], F[5] = nx * q[1] - ny * q[0];
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R c2 = c * c, FL[6], FR[6];
    normal_flux(ql, c2, n[0], n[1], FL);
    normal_flux(qr, c2, n[0], n[1], FR);
    for (int k = 0; k < 6; ++k) f[k] = R(0.5) * (FL[k] + FR[k]) - R(0.5) * c * (qr[k] - ql[k]);
  }
  // PEC wall: tangential E and normal B flip (src/local_mesh_2d_gpu_maxwell.mojo:274-292)
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    const bool wall = bc == BC_WALL, in = bc == BC_INFLOW;
    const R En = q[0] * n[0] + q[1] * n[1], Bn = q[3] * n[0] + q[4] * n[1];
    g[0] = in ? inflow[0] : (wall ? R(2) * En * n[0] - q[0] : q[0]);
    g[1] = in ? inflow[1] : (wall ? R(2) * En * n[1] - q[1] : q[1]);
    g[2] = in ? inflow[2] : (wall ? -q[2] : q[2]);
    g[3] = in ? inflow[3] : (wall ? q[3] - R(2) * Bn * n[0] : q[3]);
    g[4] = in ? inflow[4] : (wall ? q[4] - R(2) * Bn * n[1] : q[4]);
    g[5] = in ? inflow[5] : q[5];
  }
  // uniform J / M sources, src/local_mesh_2d_gpu_maxwell.mojo:133-142
  MXM_HD bool source_active() const { return true; }
  MXM_HD void source(const R *, R *s) const {
    R c2 = c * c;
    for (int d = 0; d < 3; ++d) s[d] = -(c2 * J[d]), s[3 + d] = -M[d];
  }
  // the light speed (CFL reduction)
  MXM_HD R max_speed(const R *) const { return rabs(c); }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/local_mesh_2d_gpu_advection.mojo (scalar upwind)
template <class R>
struct Advection2D {
  static constexpr int NC = 1;
  static constexpr bool kSource = false, kLimit = false;
  R vx, vy, inflow;
  MXM_HD static Advection2D from(const double *p) { return {R(p[0]), R(p[1]), R(p[2])}; }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    F[0] = vx * q[0];
    F[1] = vy * q[0];
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R vn = vx * n[0] + vy * n[1];
    f[0] = vn >= R(0) ? vn * ql[0] : vn * qr[0];
  }
  // INFLOW: user value on the inflowing half; WALL / OUTFLOW: zero
  // (src/local_mesh_2d_gpu_advection.mojo:285-298)
  MXM_HD void ghost(const R *, int bc, const R *, R *g) const { g[0] = bc == BC_INFLOW ? inflow : R(0); }
  MXM_HD void source(const R *, R *s) const { s[0] = R(0); }
  // |v| (CFL reduction)
  MXM_HD R max_speed(const R *) const { return msqrt(vx * vx + vy * vy); }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// ---------------------------------------------------------------------------
// The other trait physics of the reference's 3D solver (tets).  Each face
// flux is Rusanov, 1/2 (F_n(qL) + F_n(qR)) - 1/2 alpha (qR - qL), the
// boundary flux the same against the BC ghost.
// ---------------------------------------------------------------------------

template <class R, int NC>
MXM_HD void rusanov_combine(const R *FL, const R *FR, const R *ql, const R *qr, R alpha, R *f) {
  for (int c = 0; c < NC; ++c) f[c] = R(0.5) * (FL[c] + FR[c]) - R(0.5) * alpha * (qr[c] - ql[c]);
}
// v - 2 (v.n) n of the 3-vector at q + o (nd components take part)
template <class R>
MXM_HD void reflect3(const R *q, const R *n, int o, int nd, R *g) {
  R vn = R(0);
  for (int d = 0; d < nd; ++d) vn += q[o + d] * n[d];
  for (int d = 0; d < nd; ++d) g[o + d] = q[o + d] - R(2) * vn * n[d];
}

// src/shallow_water.mojo:31-167: (h, hu, hv), the 2D system embedded in 3D
template <class R>
struct ShallowWater3 {
  static constexpr int NC = 3;
  static constexpr bool kSource = false, kLimit = false;
  R g, hmin, inflow[3];
  MXM_HD static ShallowWater3 from(const double *p) {
    return {R(p[0]), R(p[1]), {R(p[2]), R(p[3]), R(p[4])}};
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    const R h = q[0] > hmin ? q[0] : hmin, ih = R(1) / h;
    const R hu = q[1], hv = q[2], u = hu * ih, v = hv * ih, gh2 = R(0.5) * g * h * h;
    F[0] = hu, F[1] = hu * u + gh2, F[2] = hu * v;
    F[3] = hv, F[4] = hu * v, F[5] = hv * v + gh2;
    if constexpr (ND == 3) F[6] = F[7] = F[8] = R(0);
  }
  MXM_HD R side(const R *q, R nx, R ny, R *F) const {
    const R h = q[0] > hmin ? q[0] : hmin, ih = R(1) / h;
    const R un = (q[1] * nx + q[2] * ny) * ih, gh2 = R(0.5) * g * h * h;
    F[0] = q[1] * nx + q[2] * ny, F[1] = q[1] * un + gh2 * nx, F[2] = q[2] * un + gh2 * ny;
    return rabs(un) + msqrt(g * h);
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R FL[3], FR[3];
    const R a = rmax(side(ql, n[0], n[1], FL), side(qr, n[0], n[1], FR));
    rusanov_combine<R, 3>(FL, FR, ql, qr, a, f);
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *gs) const {
    for (int c = 0; c < 3; ++c) gs[c] = bc == BC_INFLOW ? inflow[c] : q[c];
    if (bc == BC_WALL) reflect3(q, n, 1, 2, gs);
  }
  MXM_HD void source(const R *, R *s) const { s[0] = s[1] = s[2] = R(0); }
  // |u| + sqrt(g h) (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    const R h = q[0] > hmin ? q[0] : hmin;
    return msqrt(q[1] * q[1] + q[2] * q[2]) / h + msqrt(g * h);
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/maxwell.mojo:48-255: (E, B), PEC walls, uniform J and M sources
template <class R>
struct Maxwell3 {
  static constexpr int NC = 6;
  static constexpr bool kSource = true, kLimit = false;
  R c, J[3], M[3], inflow[6];
  MXM_HD static Maxwell3 from(const double *p) {
    Maxwell3 m;
    m.c = R(p[0]);
    for (int d = 0; d < 3; ++d) m.J[d] = R(p[1 + d]), m.M[d] = R(p[4 + d]);
    for (int k = 0; k < 6; ++k) m.inflow[k] = R(p[7 + k]);
    return m;
  }
  // (F.n)_E = c^2 (B x n), (F.n)_B = n x E
  MXM_HD static void side(const R *q, R c2, const R *n, R *F) {
    F[0] = c2 * (q[4] * n[2] - q[5] * n[1]);
    F[1] = c2 * (q[5] * n[0] - q[3] * n[2]);
    F[2] = c2 * (q[3] * n[1] - q[4] * n[0]);
    F[3] = n[1] * q[2] - n[2] * q[1];
    F[4] = n[2] * q[0] - n[0] * q[2];
    F[5] = n[0] * q[1] - n[1] * q[0];
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    const R c2 = c * c;
    for (int d = 0; d < ND; ++d) {
      const R e[3] = {R(d == 0), R(d == 1), R(d == 2)};
      side(q, c2, e, F + 6 * d);
    }
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R FL[6], FR[6];
    const R c2 = c * c;
    side(ql, c2, n, FL);
    side(qr, c2, n, FR);
    rusanov_combine<R, 6>(FL, FR, ql, qr, c, f);
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    for (int k = 0; k < 6; ++k) g[k] = bc == BC_INFLOW ? inflow[k] : q[k];
    if (bc == BC_WALL) {
      const R En = q[0] * n[0] + q[1] * n[1] + q[2] * n[2];
      for (int d = 0; d < 3; ++d) g[d] = R(2) * En * n[d] - q[d];
      reflect3(q, n, 3, 3, g);
    }
  }
  MXM_HD bool source_active() const { return true; }
  MXM_HD void source(const R *, R *s) const {
    const R c2 = c * c;
    for (int d = 0; d < 3; ++d) s[d] = -(c2 * J[d]), s[3 + d] = -M[d];
  }
  // the light speed (CFL reduction)
  MXM_HD R max_speed(const R *) const { return rabs(c); }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/mhd.mojo:46-367: (rho, rho u[3], E, B[3], psi), GLM cleaning with
// the damping -alpha_d psi as a source
template <class R>
struct IdealMHD3 {
  static constexpr int NC = 9;
  static constexpr bool kSource = true, kLimit = false;
  R gamma, rmin, pmin, ch, alpha, inflow[9];
  MXM_HD static IdealMHD3 from(const double *p) {
    IdealMHD3 m;
    m.gamma = R(p[0]), m.rmin = R(p[1]), m.pmin = R(p[2]), m.ch = R(p[3]), m.alpha = R(p[4]);
    for (int k = 0; k < 9; ++k) m.inflow[k] = R(p[5 + k]);
    return m;
  }
  MXM_HD R pres(const R *q, R rho, R ir) const {
    const R ke = R(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * ir;
    const R me = R(0.5) * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
    const R p = (gamma - R(1)) * (q[4] - ke - me);
    return p > pmin ? p : pmin;
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    const R rho = q[0] > rmin ? q[0] : rmin, ir = R(1) / rho;
    const R u[3] = {q[1] * ir, q[2] * ir, q[3] * ir}, B[3] = {q[5], q[6], q[7]};
    const R pt = pres(q, rho, ir) + R(0.5) * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    const R ub = u[0] * B[0] + u[1] * B[1] + u[2] * B[2], ch2 = ch * ch;
    for (int d = 0; d < ND; ++d) {
      R *Fd = F + 9 * d;
      const R m = q[1 + d];
      Fd[0] = m;
      for (int k = 0; k < 3; ++k) Fd[1 + k] = m * u[k] - B[d] * B[k];
      Fd[1 + d] += pt;
      Fd[4] = u[d] * (q[4] + pt) - B[d] * ub;
      for (int k = 0; k < 3; ++k) Fd[5 + k] = u[d] * B[k] - u[k] * B[d];
      Fd[5 + d] = q[8];
      Fd[8] = ch2 * B[d];
    }
  }
  // normal flux and |u_n| + c_f of one side (src/mhd.mojo:76-89, 152-190)
  MXM_HD R side(const R *q, const R *n, R *F) const {
    const R rho = q[0] > rmin ? q[0] : rmin, ir = R(1) / rho;
    const R u[3] = {q[1] * ir, q[2] * ir, q[3] * ir}, B[3] = {q[5], q[6], q[7]};
    const R p = pres(q, rho, ir), b2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const R pt = p + R(0.5) * b2;
    const R ub = u[0] * B[0] + u[1] * B[1] + u[2] * B[2];
    const R un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2], bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
    F[0] = rho * un;
    for (int k = 0; k < 3; ++k) F[1 + k] = rho * u[k] * un + pt * n[k] - bn * B[k];
    F[4] = un * (q[4] + pt) - bn * ub;
    for (int k = 0; k < 3; ++k) F[5 + k] = un * B[k] - u[k] * bn + q[8] * n[k];
    F[8] = ch * ch * bn;
    const R cs2 = gamma * p * ir, ca2 = b2 * ir, can2 = bn * bn * ir;
    R disc = (cs2 + ca2) * (cs2 + ca2) - R(4) * cs2 * can2;
    disc = disc < R(0) ? R(0) : disc;
    return rabs(un) + msqrt(R(0.5) * (cs2 + ca2 + msqrt(disc)));
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R FL[9], FR[9];
    const R a = rmax(rmax(side(ql, n, FL), side(qr, n, FR)), ch);
    rusanov_combine<R, 9>(FL, FR, ql, qr, a, f);
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    for (int k = 0; k < 9; ++k) g[k] = bc == BC_INFLOW ? inflow[k] : q[k];
    if (bc == BC_WALL) {
      reflect3(q, n, 1, 3, g);
      reflect3(q, n, 5, 3, g);
    }
  }
  MXM_HD bool source_active() const { return alpha != R(0); }
  MXM_HD void source(const R *q, R *s) const {
    for (int k = 0; k < 8; ++k) s[k] = R(0);
    s[8] = -alpha * q[8];
  }
  // max(|u| + the fast speed across B, c_h) (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    const R rho = q[0] > rmin ? q[0] : rmin, ir = R(1) / rho;
    const R b2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
    return nmax(msqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * ir + msqrt((gamma * pres(q, rho, ir) + b2) * ir), ch);
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

// src/two_fluid.mojo:52-513: electron (0..4) and ion (5..9) Euler fluids,
// fields E (10..12), B (13..15), psi (16); Lorentz and current sources
template <class R>
struct TwoFluid3 {
  static constexpr int NC = 17;
  static constexpr bool kSource = true, kLimit = false;
  R ge, gi, qme, qmi, ieps0, cl, ch, alpha, rmin, pmin, inflow[17];
  MXM_HD static TwoFluid3 from(const double *p) {
    TwoFluid3 t;
    t.ge = R(p[0]), t.gi = R(p[1]), t.qme = R(p[2] / p[3]), t.qmi = R(p[4] / p[5]), t.ieps0 = R(1.0 / p[6]);
    t.cl = R(p[7]), t.ch = R(p[8]), t.alpha = R(p[9]), t.rmin = R(p[10]), t.pmin = R(p[11]);
    for (int k = 0; k < 17; ++k) t.inflow[k] = R(p[12 + k]);
    return t;
  }
  MXM_HD R species_p(const R *q, R gamma, R ir) const {
    const R p = (gamma - R(1)) * (q[4] - R(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * ir);
    return p > pmin ? p : pmin;
  }
  // normal flux of one species and |u_n| + c_s (src/two_fluid.mojo:112-131)
  MXM_HD R species_side(const R *q, R gamma, const R *n, R *F) const {
    const R rho = q[0] > rmin ? q[0] : rmin, ir = R(1) / rho;
    const R p = species_p(q, gamma, ir);
    const R u[3] = {q[1] * ir, q[2] * ir, q[3] * ir}, un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    F[0] = rho * un;
    for (int k = 0; k < 3; ++k) F[1 + k] = rho * u[k] * un + p * n[k];
    F[4] = un * (q[4] + p);
    return rabs(un) + msqrt(gamma * p * ir);
  }
  MXM_HD void fields(const R *q, const R *n, R *F) const {
    const R c2 = cl * cl, ch2 = ch * ch;
    const R *E = q + 10, *B = q + 13;
    F[10] = c2 * (B[1] * n[2] - B[2] * n[1]);
    F[11] = c2 * (B[2] * n[0] - B[0] * n[2]);
    F[12] = c2 * (B[0] * n[1] - B[1] * n[0]);
    F[13] = (n[1] * E[2] - n[2] * E[1]) + q[16] * n[0];
    F[14] = (n[2] * E[0] - n[0] * E[2]) + q[16] * n[1];
    F[15] = (n[0] * E[1] - n[1] * E[0]) + q[16] * n[2];
    F[16] = ch2 * (B[0] * n[0] + B[1] * n[1] + B[2] * n[2]);
  }
  template <int ND>
  MXM_HD void flux(const R *q, R *F) const {
    for (int d = 0; d < ND; ++d) {
      const R e[3] = {R(d == 0), R(d == 1), R(d == 2)};
      R *Fd = F + 17 * d;
      for (int s0 = 0; s0 < 10; s0 += 5) {
        const R *qs = q + s0;
        const R rho = qs[0] > rmin ? qs[0] : rmin, ir = R(1) / rho;
        const R p = species_p(qs, s0 ? gi : ge, ir), m = qs[1 + d];
        Fd[s0] = m;
        for (int k = 0; k < 3; ++k) Fd[s0 + 1 + k] = m * (qs[1 + k] * ir);
        Fd[s0 + 1 + d] += p;
        Fd[s0 + 4] = (qs[1 + d] * ir) * (qs[4] + p);
      }
      fields(q, e, Fd);
    }
  }
  template <int ND = 3>
  MXM_HD void numerical_flux(const R *ql, const R *qr, const R *n, R *f) const {
    R FL[17], FR[17];
    const R ae = rmax(species_side(ql, ge, n, FL), species_side(qr, ge, n, FR));
    const R ai = rmax(species_side(ql + 5, gi, n, FL + 5), species_side(qr + 5, gi, n, FR + 5));
    fields(ql, n, FL);
    fields(qr, n, FR);
    rusanov_combine<R, 17>(FL, FR, ql, qr, rmax(rmax(ae, ai), rmax(cl, ch)), f);
  }
  MXM_HD void ghost(const R *q, int bc, const R *n, R *g) const {
    for (int k = 0; k < 17; ++k) g[k] = bc == BC_INFLOW ? inflow[k] : q[k];
    if (bc == BC_WALL) {
      reflect3(q, n, 1, 3, g);
      reflect3(q, n, 6, 3, g);
      const R En = q[10] * n[0] + q[11] * n[1] + q[12] * n[2];
      for (int d = 0; d < 3; ++d) g[10 + d] = R(2) * En * n[d] - q[10 + d];
      reflect3(q, n, 13, 3, g);
    }
  }
  MXM_HD bool source_active() const { return true; }
  MXM_HD void source(const R *q, R *s) const {
    const R re = q[0] > rmin ? q[0] : rmin, ri = q[5] > rmin ? q[5] : rmin;
    const R ire = R(1) / re, iri = R(1) / ri;
    const R ue[3] = {q[1] * ire, q[2] * ire, q[3] * ire}, ui[3] = {q[6] * iri, q[7] * iri, q[8] * iri};
    const R *E = q + 10, *B = q + 13;
    const R ve[3] = {ue[1] * B[2] - ue[2] * B[1], ue[2] * B[0] - ue[0] * B[2], ue[0] * B[1] - ue[1] * B[0]};
    const R vi[3] = {ui[1] * B[2] - ui[2] * B[1], ui[2] * B[0] - ui[0] * B[2], ui[0] * B[1] - ui[1] * B[0]};
    s[0] = s[5] = R(0);
    for (int d = 0; d < 3; ++d) {
      s[1 + d] = qme * re * (E[d] + ve[d]);
      s[6 + d] = qmi * ri * (E[d] + vi[d]);
      s[10 + d] = -(qme * re * ue[d] + qmi * ri * ui[d]) * ieps0;
      s[13 + d] = R(0);
    }
    s[4] = qme * re * (ue[0] * E[0] + ue[1] * E[1] + ue[2] * E[2]);
    s[9] = qmi * ri * (ui[0] * E[0] + ui[1] * E[1] + ui[2] * E[2]);
    s[16] = -alpha * q[16];
  }
  // max over the species of |u| + c_s, c and c_h (CFL reduction)
  MXM_HD R max_speed(const R *q) const {
    R sp[2];
    for (int k = 0; k < 2; ++k) {
      const R *qs = q + 5 * k;
      const R gm = k ? gi : ge, rho = qs[0] > rmin ? qs[0] : rmin, ir = R(1) / rho;
      sp[k] = msqrt(qs[1] * qs[1] + qs[2] * qs[2] + qs[3] * qs[3]) * ir + msqrt(gm * species_p(qs, gm, ir) * ir);
    }
    return nmax(sp[0], nmax(sp[1], rmax(rabs(cl), ch)));
  }
  template <class T>
  MXM_HD void limit(T *) const {}
};

}  // namespace mxm
