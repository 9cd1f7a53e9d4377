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
//    (18 operations);
//  * reciprocity (the propagators are symplectic) keeps (34) = -(12), so five
//    entries carry the state.
// The entries are scaled by powers of k so that a layer needs only
// X_n = 1 - c^2 / v^2 and k d, and an evanescent layer's functions by
// e^-(p+q) as var does.  The value is dltar4's times a positive factor (the
// Jacobians of the basis changes), so roots, signs and the reference's scan
// decisions are the same; a layer costs about 35 operations besides its two
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
#include "dk_physics.h"  // pow2_normalize

#define DK_HDF __host__ __device__ __forceinline__

namespace dk {

DK_HDF void layer_rec(const double4& l, const double4& up, bool top, double4* g,
                      double4* r) {
  const double irho = 1.0 / l.w;
  const double eps = top ? 1.0 : up.w * irho;
  const double dmu = top ? 0.0 : 2.0 * (up.w * up.z * up.z - l.w * l.z * l.z) * irho;
  *g = make_double4(l.x, 1.0 / (l.y * l.y), 1.0 / (l.z * l.z), eps);
  *r = make_double4(dmu, 0.0, 0.0, 0.0);
}

// One wave type of a layer for X = k^2 - kx^2: C = cosh(p), S = sinh(p)/sqrt(X),
// R = sqrt(X) sinh(p), p = sqrt(X) d, all scaled by exp(-p) (X > 0), or
// cos, sin/sqrt(-X), -sqrt(-X) sin (X < 0).  Returns the scale, *ex = p or 0.
// (exp(-2p) is kept where the reference's var drops it for p >= 16: a change
// of F below 1e-13 relative.)
DK_HDF double hyp_fwd(double X, double d, double* C, double* S, double* R, double* ex) {
  // X = 0 (the trial velocity equal to the layer's) is the limit of the
  // oscillatory form: taken at X = -1e-300 it gives C = 1, S = d, R = 0 to
  // rounding, without a branch of its own
  if (X == 0.0) X = -1.0e-300;
  const double ax = fabs(X);
  const double y = rsqrt_nr(ax);       // 1/sqrt|X|
  const double ra = ax * y;            // sqrt|X|
  const double p = ra * d;
  if (X > 0.0) {
    const double s = exp_neg(p);
    const double fac = s * s;
    const double sh = 0.5 - 0.5 * fac;
    *C = 0.5 + 0.5 * fac;
    *S = sh * y;
    *R = ra * sh;
    *ex = p;
    return s;
  }
  double sp, cp;
  sincos_phase(p, &sp, &cp);
  *C = cp;
  *S = sp * y;
  *R = -ra * sp;
  *ex = 0.0;
  return 1.0;
}

// State e = (S11, S12, S22, Y12, Y21): S11 = (13), S12 = k (12), S22 = -k^2 (24),
// Y12 = -k (14), Y21 = k (23) of the compound row vector, up to a common factor.
// Propagation through a layer with the record g = (d, 1/a^2, 1/b^2, -).
DK_HDF void fdm_layer(double c2, double k, const double4& g, double e[5]) {
  const double kd = k * g.x;
  double Ca, Sa, Ra, Cb, Sb, Rb, pa, pb;
  const double sa = hyp_fwd(fma(-c2, g.y, 1.0), kd, &Ca, &Sa, &Ra, &pa);
  const double sb = hyp_fwd(fma(-c2, g.z, 1.0), kd, &Cb, &Sb, &Rb, &pb);
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

// The interface above a layer: eps, dmu of its record, ic2 = 1/c^2.
DK_HDF void fdm_iface(double ic2, double eps, double dmu, double e[5]) {
  const double eta = dmu * ic2;
  const double a = eps - eta, am1 = a - 1.0, b = 1.0 + eta;
  const double s11 = e[0], s12 = e[1], s22 = e[2];
  const double y1 = fma(a, s11, am1 * s12), y2 = fma(a, s12, am1 * s22);
  const double z1 = fma(eta, s11, b * s12), z2 = fma(eta, s12, b * s22);
  e[0] = fma(a, y1, am1 * y2);
  e[1] = fma(eta, y1, b * y2);
  e[2] = fma(eta, z1, b * z2);
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
  fdm_iface(ic2, gH.w, mdl[2 * (L - 1) + 1].x, e);
  int E = 0;
  // two layers per step; one normalisation per two steps keeps the state far
  // inside the double range.  The next pair's records are read from LDS one
  // step ahead (clamped indices), so that the read latency overlaps this
  // pair's work.
  int m = L - 2;
  auto cl = [&](int i) { return i < 0 ? 0 : (i > L - 1 ? L - 1 : i); };  // reads stay in range
  int ia = cl(m), ib = cl(m - 1);
  double4 ga = mdl[2 * ia], gb = mdl[2 * ib];
  double da = mdl[2 * ia + 1].x, db = mdl[2 * ib + 1].x;
  bool norm = false;
  for (; m >= 1; m -= 2) {
    const double4 g0 = ga, g1 = gb;
    const double d0 = da, d1 = db;
    ia = cl(m - 2);
    ib = cl(m - 3);
    ga = mdl[2 * ia];
    gb = mdl[2 * ib];
    da = mdl[2 * ia + 1].x;
    db = mdl[2 * ib + 1].x;
    fdm_layer(c2, k, g0, e);
    fdm_iface(ic2, g0.w, d0, e);
    fdm_layer(c2, k, g1, e);
    fdm_iface(ic2, g1.w, d1, e);  // the identity above the top layer
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

}  // namespace dk
