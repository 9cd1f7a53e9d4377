// Value of the Rayleigh secular function for the root search, written once
// for the device (k_forward) and a host build that pins it against the
// oracle's dltar4 in the CPU tests.
//
// The root search only needs the sign and value of F(c, omega), the Rayleigh
// determinant that dltar4 forms from the Dunkin compound matrices of the layers
// (surfdisp96.f:767-865 with var and dnka).  Here the same determinant is
// propagated in each layer's P-SV potential basis (phi, phi', psi, psi'), in
// which the compound matrices factor:
//  * inside a layer the potentials propagate block-diagonally, P_a (+) P_b with
//    P = [[C, S/r], [r S, C]] of unit determinant, so the compound keeps its
//    (12) and (34) entries and maps the 2x2 array Y of the mixed entries
//    (13) (14) / (23) (24) to P_a^T Y P_b (16 multiply-adds);
//  * at an interface the change of basis M_lo^-1 M_up is two 2x2 blocks with
//    the same four entries; its compound maps the symmetric array
//    S = [[(13), (12)], [(12), (24)]] to N S N^T with
//    N = [[a, a - 1], [eta, 1 + eta]], a = eps - eta, eps = rho_up / rho_lo,
//    eta = 2 (mu_up - mu_lo) / (rho_lo c^2), and scales (14), (23) by eps
//    (13 operations in the rank-one form of fdm_iface);
//  * reciprocity (the propagators are symplectic) keeps (34) = -(12), so five
//    entries carry the state.
// The entries are scaled by powers of k so that a layer needs only
// X_n = 1 - c^2 / v^2 and k d, and an evanescent layer's functions by
// e^-(p+q) as var does.  The value is dltar4's times a positive factor (the
// Jacobians of the basis changes), so roots, signs and the reference's scan
// decisions are the same; a layer costs about 30 operations besides its two
// wave functions, against about 95 for building and applying the Dunkin
// matrix.  (This is the factorisation of the fast delta matrix, Buchen and
// Ben-Hador, Geophys. J. Int. 124 (1996) 869-887.)
//
// Layer record of layer m (two double4, staged in LDS by the root search):
// (d, 1/a^2, 1/b^2, eps_m) and (dmu_m, -, -, -), with eps_m = rho_{m-1}/rho_m and
// dmu_m = 2 (mu_{m-1} - mu_m) / rho_m for the interface above layer m (the
// identity, eps = 1 and dmu = 0, above the top layer).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "dk_fastmath.h"

#define DK_HDF __host__ __device__ __forceinline__

namespace dk {

// Exact power-of-two normalisation of a 5-entry state: scales e by 2^-k with
// k the exponent of max|e| and returns k (0 for a zero or non-finite state).
DK_HD int pow2_normalize(double e[5]) {
  double m = fabs(e[0]);
#pragma unroll
  for (int i = 1; i < 5; i++) m = fmax(m, fabs(e[i]));
  if (!(m > 1e-300) || !isfinite(m)) return 0;
  int k;
  frexp(m, &k);
#pragma unroll
  for (int i = 0; i < 5; i++) e[i] = ldexp(e[i], -k);
  return k;
}

DK_HDF void layer_rec(const double4& l, const double4& up, bool top, double4* g,
                      double4* r) {
  const double irho = 1.0 / l.w;
  const double eps = top ? 1.0 : up.w * irho;
  const double dmu = top ? 0.0 : 2.0 * (up.w * up.z * up.z - l.w * l.z * l.z) * irho;
  *g = make_double4(l.x, 1.0 / (l.y * l.y), 1.0 / (l.z * l.z), eps);
  *r = make_double4(dmu, top ? 0.0 : up.w * irho - 1.0, 0.0, 0.0);
}

// One wave type of a layer for X = k^2 - kx^2 (the root search's evaluation,
// with the shorter exp and sincos): C = cosh(p), S = sinh(p)/sqrt(X),
// R = sqrt(X) sinh(p), p = sqrt(X) d, all scaled by exp(-p) (X > 0), or
// cos, sin/sqrt(-X), -sqrt(-X) sin (X < 0).  Returns the scale, *ex = p or 0.
// (exp(-2p) is kept where the reference's var drops it for p >= 16: a change
// of F below 1e-13 relative.)
DK_HDF double hyp_fwd(double X, double d, double* C, double* S, double* R, double* ex) {
  // X = 0 (the trial velocity equal to the layer's) is the limit of the
  // oscillatory form: with |X| + 1e-300 (which changes no X != 0: |X| is at
  // least the rounding granularity of 1 - c^2/v^2) it gives C = 1, S = d,
  // R = 0 to rounding, without a select of its own
  const double ax = fabs(X) + 1.0e-300;
  const double y = rsqrt_nr(ax);       // 1/sqrt|X|
  const double ra = ax * y;            // sqrt|X|
  const double p = ra * d;
  if (X > 0.0) {
    const double s = exp_neg_rs(p);
    const double fac = s * s;
    const double sh = 0.5 - 0.5 * fac;
    *C = 0.5 + 0.5 * fac;
    *S = sh * y;
    *R = ra * sh;
    *ex = p;
    return s;
  }
  double sp, cp;
  sincos_phase_rs(p, &sp, &cp);
  *C = cp;
  *S = sp * y;
  *R = -ra * sp;
  *ex = 0.0;
  return 1.0;
}

// State e = (S11, S12, S22, Y12, Y21): S11 = (13), S12 = k (12), S22 = -k^2 (24),
// Y12 = -k (14), Y21 = k (23) of the compound row vector, up to a common factor.
// Propagation through a layer with the record g = (d, 1/a^2, 1/b^2, -).
// The evanescent form of hyp_fwd (X > 0), without its branch.
DK_HDF double hyp_ev(double X, double d, double* C, double* S, double* R) {
  const double y = rsqrt_nr(X);
  const double ra = X * y;
  const double s = exp_neg_rs(ra * d);
  const double fac = s * s;
  const double sh = 0.5 - 0.5 * fac;
  *C = 0.5 + 0.5 * fac;
  *S = sh * y;
  *R = ra * sh;
  return s;
}

DK_HDF void fdm_layer(double c2, double k, const double4& g, double e[5]) {
  const double kd = k * g.x;
  const double XP = fma(-c2, g.y, 1.0), XS = fma(-c2, g.z, 1.0);
  double Ca, Sa, Ra, Cb, Sb, Rb, sa, sb;
  // Both wave types evanescent on every lane of the wave (the common case
  // below the top layers): one wave-uniform branch instead of two divergent
  // ones, whose exec-mask bookkeeping is scalar work on the wave's own
  // timeline.
#if defined(__HIP_DEVICE_COMPILE__)
  const bool all_ev = __all(XP > 0.0 && XS > 0.0);
#else
  const bool all_ev = XP > 0.0 && XS > 0.0;
#endif
  if (all_ev) {
    sa = hyp_ev(XP, kd, &Ca, &Sa, &Ra);
    sb = hyp_ev(XS, kd, &Cb, &Sb, &Rb);
  } else {
    double pa, pb;
    sa = hyp_fwd(XP, kd, &Ca, &Sa, &Ra, &pa);
    sb = hyp_fwd(XS, kd, &Cb, &Sb, &Rb, &pb);
  }
  const double s11 = e[0], s22 = e[2], y12 = e[3], y21 = e[4];
  // P_a^T Y, P_a^T = [[Ca, Ra], [Sa, Ca]] (k-scaled)
  const double t11 = fma(Ca, s11, Ra * y21), t12 = fma(Ca, y12, Ra * s22);
  const double t21 = fma(Sa, s11, Ca * y21), t22 = fma(Sa, y12, Ca * s22);
  // ... P_b, P_b = [[Cb, -Sb], [-Rb, Cb]] (k-scaled)
  e[0] = fma(t11, Cb, -t12 * Rb);
  e[3] = fma(t12, Cb, -t11 * Sb);
  e[4] = fma(t21, Cb, -t22 * Rb);
  e[2] = fma(t22, Cb, -t21 * Sb);
  e[1] *= sa * sb;  // (12) passes the layer unchanged: only the layer's scale
}

// The interface above a layer: eps, eps - 1 and dmu of its record, ic2 = 1/c^2.
// N = [[a, a - 1], [eta, 1 + eta]] is I + n (1, 1) with n = (a - 1, eta), so
// N S N^T = S + n q^T + q n^T + (q1 + q2) n n^T with q = S (1, 1)^T.
DK_HDF void fdm_iface(double ic2, double eps, double epsm1, double dmu, double e[5]) {
  const double n2 = dmu * ic2;  // eta
  const double n1 = epsm1 - n2;  // a - 1
  const double q1 = e[0] + e[1], q2 = e[1] + e[2], sg = q1 + q2;
  const double t1 = fma(sg, n1, q1), t2 = fma(sg, n2, q2);
  const double s12 = fma(n1, t2, fma(n2, q1, e[1]));
  e[0] = fma(n1, q1 + t1, e[0]);
  e[2] = fma(n2, q2 + t2, e[2]);
  e[1] = s12;
  e[3] *= eps;
  e[4] *= eps;
}

// F(c, omega) as mantissa and exponent over the LDS-staged records of L layers
// (half-space last).
DK_HDF double secular_fwd(double c, double omega, int L, const double4* mdl, int* ex) {
  if (omega < 1.0e-4) omega = 1.0e-4;
  const double k = omega / c, c2 = c * c, ic2 = 1.0 / c2;
  // the half-space's decaying P and S potentials, then the interface above it
  const double4 gH = mdl[2 * (L - 1)];
  const double ran = sqrt(fabs(fma(-c2, gH.y, 1.0)));
  const double rbn = sqrt(fabs(fma(-c2, gH.z, 1.0)));
  double e[5] = {ran * rbn, 0.0, -1.0, -ran, rbn};
  {
    const double4 rH = mdl[2 * (L - 1) + 1];
    fdm_iface(ic2, gH.w, rH.y, rH.x, e);
  }
  int E = 0;
  // two layers per step; one normalisation per two steps keeps the state far
  // inside the double range.  The next pair's records are read from LDS one
  // step ahead (clamped indices), so that the read latency overlaps this
  // pair's work.
  int m = L - 2;
  auto cl = [&](int i) { return i < 0 ? 0 : (i > L - 1 ? L - 1 : i); };  // reads stay in range
  int ia = cl(m), ib = cl(m - 1);
  double4 ga = mdl[2 * ia], gb = mdl[2 * ib];
  double4 ra = mdl[2 * ia + 1], rb = mdl[2 * ib + 1];
  bool norm = false;
  for (; m >= 1; m -= 2) {
    const double4 g0 = ga, g1 = gb, r0 = ra, r1 = rb;
    ia = cl(m - 2);
    ib = cl(m - 3);
    ga = mdl[2 * ia];
    gb = mdl[2 * ib];
    ra = mdl[2 * ia + 1];
    rb = mdl[2 * ib + 1];
    fdm_layer(c2, k, g0, e);
    fdm_iface(ic2, g0.w, r0.y, r0.x, e);
    fdm_layer(c2, k, g1, e);
    fdm_iface(ic2, g1.w, r1.y, r1.x, e);  // the identity above the top layer
    if (norm) E += pow2_normalize(e);
    norm = !norm;
  }
  if (m == 0) fdm_layer(c2, k, mdl[0], e);
  E += pow2_normalize(e);
  // free surface of the top layer: F = -(4 S11 + 4 t S12 + t^2 S22),
  // t = 2 - c^2 / b_0^2
  const double t = fma(-c2, mdl[0].z, 2.0);
  *ex = E;
  return -fma(t, fma(t, e[2], 4.0 * e[1]), 4.0 * e[0]);
}

// ------------------------------------------------------------------ derivatives
// The kernel pass and the VJP differentiate the same function: dc/dm =
// -(dF/dm)/(dF/dc) at the root, where F's positive factor drops out (F = 0).
// A bottom-up sweep carries the state and its tangent in c; a top-down
// adjoint sweep carries dF/d(state) and, per layer, the partials in the
// flattened (a, b, rho, d).  The building blocks below are the forward step,
// its tangent and its adjoint for each part of the recursion; the partials of
// a wave type's functions follow from d/dp and d/dr (p = r kd, r = sqrt|X|).

// One wave type of a layer: the values of hyp_fwd with their p-derivatives
// dC/dp = Cp, dS/dp = Q y, dR/dp = 2 h Q r, ds/dp = sp (dS/dr = -S y and
// dR/dr = R y in both branches), r, y = 1/r and h = +-1/2 (dr/dX = h y).
struct WaveFn {
  double C, S, R, s;
  double Cp, Q, sp;
  double r, y, h;
};

DK_HDF void wave_fn(double X, double kd, WaveFn& w) {
  if (X == 0.0) X = -1.0e-300;  // the oscillatory limit, as in hyp_fwd
  const double ax = fabs(X);
  const double y = rsqrt_nr(ax);
  const double r = ax * y;
  const double p = r * kd;
  w.r = r;
  w.y = y;
  if (X > 0.0) {
    const double s = exp_neg(p), f = s * s, sh = 0.5 - 0.5 * f;
    w.C = 0.5 + 0.5 * f;  // cosh(p) e^-p
    w.S = sh * y;         // sinh(p) e^-p / r
    w.R = r * sh;         // r sinh(p) e^-p
    w.s = s;
    w.Cp = -f;
    w.Q = f;
    w.sp = -s;
    w.h = 0.5;
  } else {
    double sp, cp;
    sincos_phase(p, &sp, &cp);
    w.C = cp;
    w.S = sp * y;
    w.R = -r * sp;
    w.s = 1.0;
    w.Cp = -sp;
    w.Q = cp;  // dS/dp = cos(p)/r, dR/dp = -r cos(p)
    w.sp = 0.0;
    w.h = -0.5;
  }
}

// Tangent of a wave type's functions for dX, dkd: (dC, dS, dR, ds).
DK_HDF void wave_tan(const WaveFn& w, double kd, double dX, double dkd, double d[4]) {
  const double dr = w.h * w.y * dX;
  const double dp = fma(kd, dr, w.r * dkd);
  d[0] = w.Cp * dp;
  d[1] = w.y * fma(w.Q, dp, -w.S * dr);
  d[2] = fma(2.0 * w.h * w.Q * w.r, dp, w.R * w.y * dr);
  d[3] = w.sp * dp;
}

// Adjoint of a wave type's functions: from (Cb, Sb, Rb, sb) = dF/d(C, S, R, s)
// to dF/dX (*Xb) and dF/dkd (*kdb).
DK_HDF void wave_adj(const WaveFn& w, double kd, const double b[4], double* Xb, double* kdb) {
  const double qb = w.Q * fma(b[1], w.y, 2.0 * w.h * w.r * b[2]);
  const double pb = fma(b[0], w.Cp, fma(b[3], w.sp, qb));
  const double rb = fma(pb, kd, w.y * fma(b[2], w.R, -b[1] * w.S));
  *Xb = rb * w.h * w.y;
  *kdb = pb * w.r;
}

// Layer step u = L(v) with the intermediates t = P_a^T Y.
DK_HDF void fdm_layer_fn(const WaveFn& A, const WaveFn& B, const double v[5], double t[4],
                         double u[5]) {
  t[0] = fma(A.C, v[0], A.R * v[4]);
  t[1] = fma(A.C, v[3], A.R * v[2]);
  t[2] = fma(A.S, v[0], A.C * v[4]);
  t[3] = fma(A.S, v[3], A.C * v[2]);
  u[0] = fma(t[0], B.C, -t[1] * B.R);
  u[3] = fma(t[1], B.C, -t[0] * B.S);
  u[4] = fma(t[2], B.C, -t[3] * B.R);
  u[2] = fma(t[3], B.C, -t[2] * B.S);
  u[1] = v[1] * (A.s * B.s);
}

// Its tangent: dA, dB = (dC, dS, dR, ds) of the two wave types.
DK_HDF void fdm_layer_tan(const WaveFn& A, const WaveFn& B, const double dA[4], const double dB[4],
                          const double v[5], const double dv[5], const double t[4],
                          double du[5]) {
  const double dt0 = fma(dA[0], v[0], fma(A.C, dv[0], fma(dA[2], v[4], A.R * dv[4])));
  const double dt1 = fma(dA[0], v[3], fma(A.C, dv[3], fma(dA[2], v[2], A.R * dv[2])));
  const double dt2 = fma(dA[1], v[0], fma(A.S, dv[0], fma(dA[0], v[4], A.C * dv[4])));
  const double dt3 = fma(dA[1], v[3], fma(A.S, dv[3], fma(dA[0], v[2], A.C * dv[2])));
  du[0] = fma(dt0, B.C, fma(t[0], dB[0], -fma(dt1, B.R, t[1] * dB[2])));
  du[3] = fma(dt1, B.C, fma(t[1], dB[0], -fma(dt0, B.S, t[0] * dB[1])));
  du[4] = fma(dt2, B.C, fma(t[2], dB[0], -fma(dt3, B.R, t[3] * dB[2])));
  du[2] = fma(dt3, B.C, fma(t[3], dB[0], -fma(dt2, B.S, t[2] * dB[1])));
  du[1] = fma(dv[1], A.s * B.s, v[1] * fma(dA[3], B.s, A.s * dB[3]));
}

// Its adjoint: ub = dF/du -> vb = dF/dv, and ab, bb = dF/d(C, S, R, s) of the
// two wave types.
DK_HDF void fdm_layer_adj(const WaveFn& A, const WaveFn& B, const double v[5], const double t[4],
                          const double ub[5], double vb[5], double ab[4], double bb[4]) {
  const double tb0 = fma(ub[0], B.C, -ub[3] * B.S);
  const double tb1 = fma(ub[3], B.C, -ub[0] * B.R);
  const double tb2 = fma(ub[4], B.C, -ub[2] * B.S);
  const double tb3 = fma(ub[2], B.C, -ub[4] * B.R);
  bb[0] = fma(ub[0], t[0], fma(ub[3], t[1], fma(ub[4], t[2], ub[2] * t[3])));
  bb[1] = -fma(ub[3], t[0], ub[2] * t[2]);
  bb[2] = -fma(ub[0], t[1], ub[4] * t[3]);
  const double us = ub[1] * v[1];
  bb[3] = us * A.s;
  ab[3] = us * B.s;
  ab[0] = fma(tb0, v[0], fma(tb1, v[3], fma(tb2, v[4], tb3 * v[2])));
  ab[1] = fma(tb2, v[0], tb3 * v[3]);
  ab[2] = fma(tb0, v[4], tb1 * v[2]);
  vb[0] = fma(tb0, A.C, tb2 * A.S);
  vb[4] = fma(tb0, A.R, tb2 * A.C);
  vb[3] = fma(tb1, A.C, tb3 * A.S);
  vb[2] = fma(tb1, A.R, tb3 * A.C);
  vb[1] = ub[1] * (A.s * B.s);
}

// Interface step w = I(u) with the intermediates yz = (y1, y2, z1, z2).
DK_HDF void fdm_iface_fn(double eps, double eta, const double u[5], double yz[4], double w[5]) {
  const double a = eps - eta, am1 = a - 1.0, b = 1.0 + eta;
  yz[0] = fma(a, u[0], am1 * u[1]);
  yz[1] = fma(a, u[1], am1 * u[2]);
  yz[2] = fma(eta, u[0], b * u[1]);
  yz[3] = fma(eta, u[1], b * u[2]);
  w[0] = fma(a, yz[0], am1 * yz[1]);
  w[1] = fma(eta, yz[0], b * yz[1]);
  w[2] = fma(eta, yz[2], b * yz[3]);
  w[3] = eps * u[3];
  w[4] = eps * u[4];
}

// Its tangent for d(eta) (eps does not depend on c).
DK_HDF void fdm_iface_tan(double eps, double eta, double deta, const double u[5],
                          const double du[5], const double yz[4], double dw[5]) {
  const double a = eps - eta, am1 = a - 1.0, b = 1.0 + eta;
  const double u01 = u[0] + u[1], u12 = u[1] + u[2];
  const double dy1 = fma(-deta, u01, fma(a, du[0], am1 * du[1]));
  const double dy2 = fma(-deta, u12, fma(a, du[1], am1 * du[2]));
  const double dz1 = fma(deta, u01, fma(eta, du[0], b * du[1]));
  const double dz2 = fma(deta, u12, fma(eta, du[1], b * du[2]));
  dw[0] = fma(-deta, yz[0] + yz[1], fma(a, dy1, am1 * dy2));
  dw[1] = fma(deta, yz[0] + yz[1], fma(eta, dy1, b * dy2));
  dw[2] = fma(deta, yz[2] + yz[3], fma(eta, dz1, b * dz2));
  dw[3] = eps * du[3];
  dw[4] = eps * du[4];
}

// Its adjoint: wb -> ub, with dF/deps (*epsb) and dF/deta (*etab).
DK_HDF void fdm_iface_adj(double eps, double eta, const double u[5], const double yz[4],
                          const double wb[5], double ub[5], double* epsb, double* etab) {
  const double a = eps - eta, am1 = a - 1.0, b = 1.0 + eta;
  const double yb1 = fma(wb[0], a, wb[1] * eta);
  const double yb2 = fma(wb[0], am1, wb[1] * b);
  const double zb1 = wb[2] * eta;
  const double zb2 = wb[2] * b;
  ub[0] = fma(yb1, a, zb1 * eta);
  ub[1] = fma(yb1, am1, fma(yb2, a, fma(zb1, b, zb2 * eta)));
  ub[2] = fma(yb2, am1, zb2 * b);
  ub[3] = wb[3] * eps;
  ub[4] = wb[4] * eps;
  const double u01 = u[0] + u[1], u12 = u[1] + u[2];
  const double abar = fma(wb[0], yz[0] + yz[1], fma(yb1, u01, yb2 * u12));
  *etab = fma(wb[1], yz[0] + yz[1], fma(wb[2], yz[2] + yz[3], fma(zb1, u01, zb2 * u12))) - abar;
  *epsb = fma(wb[3], u[3], fma(wb[4], u[4], abar));
}

// Interface constants of the flattened layers up (above) and lo (below),
// double4 (d, a, b, rho): eps = rho_up / rho_lo, dmu = 2 (mu_up - mu_lo) / rho_lo.
DK_HDF void iface_consts(const double4& up, const double4& lo, double* eps, double* dmu) {
  const double irho = 1.0 / lo.w;
  *eps = up.w * irho;
  *dmu = 2.0 * (up.w * up.z * up.z - lo.w * lo.z * lo.z) * irho;
}

// dF/deps, dF/deta of an interface -> partials in (rho, b) of the layers above
// and below it (ic2 = 1/c^2).
DK_HDF void iface_chain(const double4& up, const double4& lo, double eps, double dmu, double ic2,
                        double epsb, double etab, double* rho_up, double* b_up, double* rho_lo,
                        double* b_lo) {
  const double irho = 1.0 / lo.w;
  const double mb = etab * ic2;  // dF/d(dmu)
  *rho_up = fma(epsb, irho, mb * 2.0 * up.z * up.z * irho);
  *b_up = mb * 4.0 * up.w * up.z * irho;
  *rho_lo = -fma(epsb, eps, mb * fma(2.0 * lo.z, lo.z, dmu)) * irho;
  *b_lo = -mb * 4.0 * lo.z;
}

// Half-space start state for (a_H, b_H), its tangent in c and its adjoint.
DK_HDF void fdm_start(double c2, const double4& h, double v[5], double* ran, double* rbn,
                      double* sa, double* sb) {
  const double xa = fma(-c2, 1.0 / (h.y * h.y), 1.0), xb = fma(-c2, 1.0 / (h.z * h.z), 1.0);
  *ran = sqrt(fabs(xa));
  *rbn = sqrt(fabs(xb));
  *sa = xa < 0.0 ? -1.0 : 1.0;
  *sb = xb < 0.0 ? -1.0 : 1.0;
  v[0] = *ran * *rbn;
  v[1] = 0.0;
  v[2] = -1.0;
  v[3] = -*ran;
  v[4] = *rbn;
}
// d sqrt|1 - c^2/v^2| for d(c^2/v^2) = q: -sgn q / (2 r) (0 where r = 0)
DK_HDF double start_dr(double r, double sg, double q) { return r > 0.0 ? -sg * q / (2.0 * r) : 0.0; }

// ------------------------------------------------------------------ host reference
// Partials of F at (c, omega) in the flattened (a, b, rho, d) of each layer,
// scaled to kernels -(dF/dm)/(dF/dc): the same building blocks in plain loops
// with a full tape (host only; the device passes interleave loads and stores).
// lay: L double4 (d, a, b, rho), half-space last; out: 4 L (a, b, rho, d).
// Returns dF/dc as mantissa with *ex.
inline double fdm_kernels_ref(double c, double omega, int L, const double4* lay, double* out,
                              int* ex) {
  if (omega < 1.0e-4) omega = 1.0e-4;
  const double k = omega / c, c2 = c * c, ic2 = 1.0 / c2, dk = -k / c;
  const int H = L - 1;
  double V[1024][5];
  int EV[1024];
  double v[5], dv[5], ran, rbn, sgA, sgB;
  fdm_start(c2, lay[H], v, &ran, &rbn, &sgA, &sgB);
  {
    const double qa = 2.0 * c / (lay[H].y * lay[H].y), qb = 2.0 * c / (lay[H].z * lay[H].z);
    const double dra = start_dr(ran, sgA, qa), drb = start_dr(rbn, sgB, qb);
    dv[0] = fma(dra, rbn, ran * drb);
    dv[1] = 0.0;
    dv[2] = 0.0;
    dv[3] = -dra;
    dv[4] = drb;
  }
  int E = 0;
  auto iface_up = [&](int m) {  // the interface above layer m (m >= 1)
    double eps, dmu, yz[4], w[5], dw[5];
    iface_consts(lay[m - 1], lay[m], &eps, &dmu);
    const double eta = dmu * ic2, deta = -2.0 * eta / c;
    fdm_iface_fn(eps, eta, v, yz, w);
    fdm_iface_tan(eps, eta, deta, v, dv, yz, dw);
    for (int i = 0; i < 5; i++) {
      v[i] = w[i];
      dv[i] = dw[i];
    }
  };
  if (H >= 1) iface_up(H);
  for (int m = H - 1; m >= 0; m--) {
    for (int i = 0; i < 5; i++) V[m][i] = v[i];
    EV[m] = E;
    const double4 l = lay[m];
    const double kd = k * l.x, dkd = dk * l.x;
    WaveFn A, B;
    wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
    wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, B);
    double dA[4], dB[4], t[4], u[5], du[5];
    wave_tan(A, kd, -2.0 * c / (l.y * l.y), dkd, dA);
    wave_tan(B, kd, -2.0 * c / (l.z * l.z), dkd, dB);
    fdm_layer_fn(A, B, v, t, u);
    fdm_layer_tan(A, B, dA, dB, v, dv, t, du);
    for (int i = 0; i < 5; i++) {
      v[i] = u[i];
      dv[i] = du[i];
    }
    if (m >= 1) iface_up(m);
    double mx = 0.0;
    for (int i = 0; i < 5; i++) mx = fmax(mx, fabs(v[i]));
    int kk = 0;
    if (mx > 1e-300 && isfinite(mx)) frexp(mx, &kk);
    for (int i = 0; i < 5; i++) {
      v[i] = ldexp(v[i], -kk);
      dv[i] = ldexp(dv[i], -kk);
    }
    E += kk;
  }
  const double ib0 = 1.0 / (lay[0].z * lay[0].z);
  const double t = fma(-c2, ib0, 2.0), dt = -2.0 * c * ib0;
  const double tb = -fma(2.0 * t, v[2], 4.0 * v[1]);  // dF/dt
  const double dF = -fma(t, fma(t, dv[2], 4.0 * dv[1]), 4.0 * dv[0]) + tb * dt;
  const int EF = E;
  *ex = EF;
  const double sc = dF != 0.0 ? -1.0 / dF : 0.0;
  for (int i = 0; i < 4 * L; i++) out[i] = 0.0;
  auto add = [&](int m, int p, double val, int e) { out[4 * m + p] += ldexp(val, e - EF) * sc; };
  add(0, 1, tb * 2.0 * c2 * ib0 / lay[0].z, EF);
  double g[5] = {-4.0, -4.0 * t, -t * t, 0.0, 0.0};
  int Eg = 0;
  for (int m = 0; m <= H - 1; m++) {
    const double4 l = lay[m];
    const double kd = k * l.x;
    WaveFn A, B;
    wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
    wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, B);
    double t4[4], u[5], ub[5], vb[5], ab[4], bb[4];
    fdm_layer_fn(A, B, V[m], t4, u);
    const int e = EV[m] + Eg;
    if (m >= 1) {
      double eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
      iface_consts(lay[m - 1], l, &eps, &dmu);
      fdm_iface_fn(eps, dmu * ic2, u, yz, w);
      fdm_iface_adj(eps, dmu * ic2, u, yz, g, ub, &epsb, &etab);
      iface_chain(lay[m - 1], l, eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
      add(m - 1, 2, ru, e);
      add(m - 1, 1, bu, e);
      add(m, 2, rl, e);
      add(m, 1, bl, e);
    } else {
      for (int i = 0; i < 5; i++) ub[i] = g[i];
    }
    fdm_layer_adj(A, B, V[m], t4, ub, vb, ab, bb);
    double Xa, kda, Xb, kdb;
    wave_adj(A, kd, ab, &Xa, &kda);
    wave_adj(B, kd, bb, &Xb, &kdb);
    add(m, 0, Xa * 2.0 * c2 / (l.y * l.y * l.y), e);
    add(m, 1, Xb * 2.0 * c2 / (l.z * l.z * l.z), e);
    add(m, 3, (kda + kdb) * k, e);
    for (int i = 0; i < 5; i++) g[i] = vb[i];
    Eg += pow2_normalize(g);
  }
  // the half-space: the interface above it, then the start state
  double ub[5];
  if (H >= 1) {
    double v0[5], r1, r2, s1, s2, eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
    fdm_start(c2, lay[H], v0, &r1, &r2, &s1, &s2);
    iface_consts(lay[H - 1], lay[H], &eps, &dmu);
    fdm_iface_fn(eps, dmu * ic2, v0, yz, w);
    fdm_iface_adj(eps, dmu * ic2, v0, yz, g, ub, &epsb, &etab);
    iface_chain(lay[H - 1], lay[H], eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
    add(H - 1, 2, ru, Eg);
    add(H - 1, 1, bu, Eg);
    add(H, 2, rl, Eg);
    add(H, 1, bl, Eg);
  } else {
    for (int i = 0; i < 5; i++) ub[i] = g[i];
  }
  // start: v = (ran rbn, 0, -1, -ran, rbn)
  const double rab = fma(ub[0], rbn, -ub[3]), rbb = fma(ub[0], ran, ub[4]);
  const double4 h = lay[H];
  add(H, 0, rab * start_dr(ran, sgA, -2.0 * c2 / (h.y * h.y * h.y)), Eg);
  add(H, 1, rbb * start_dr(rbn, sgB, -2.0 * c2 / (h.z * h.z * h.z)), Eg);
  return dF;
}

}  // namespace dk
