// Rayleigh-wave physics of the depth-kernel stage, written once for the
// device (and a host build used by CPU self-checks of the derivative code).
//
// Secular function: Dunkin compound-matrix recursion from the half-space up
// (reference: code/src_inv_iso_joint/surfdisp96.f:767-865 dltar4, :868-985 var,
// :1018-1062 dnka). Differences from the reference, by design:
//   * normalisation of the 5-vector after each layer is an exact power of two
//     (frexp/ldexp) instead of a division by max|e| (surfdisp96.f:989-1014),
//     and the exponents are accumulated, so the returned value m*2^E is the
//     un-normalised, exp(-exa)-scaled determinant: continuous in c and with the
//     reference's sign;
//   * every function is templated on the number type, so the same code
//     evaluated on Dual<N> yields exact partial derivatives of each layer's
//     compound matrix with respect to (alpha, beta, rho, c).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dk_fastmath.h"

namespace dk {

// ----------------------------------------------------------------- dual numbers
template <int N>
struct Dual {
  double v;
  double g[N];
};

template <int N>
DK_HD Dual<N> dconst(double v) {
  Dual<N> r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = 0.0;
  return r;
}
template <int N>
DK_HD Dual<N> dvar(double v, int idx) {
  Dual<N> r = dconst<N>(v);
  r.g[idx] = 1.0;
  return r;
}
template <int N>
DK_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] + b.g[i];
  return r;
}
template <int N>
DK_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] - b.g[i];
  return r;
}
template <int N>
DK_HD Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = -a.g[i];
  return r;
}
template <int N>
DK_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
template <int N>
DK_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  double ib = 1.0 / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = (a.g[i] - r.v * b.g[i]) * ib;
  return r;
}
template <int N>
DK_HD Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v += b;
  return r;
}
template <int N>
DK_HD Dual<N> operator+(double b, const Dual<N>& a) {
  return a + b;
}
template <int N>
DK_HD Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v -= b;
  return r;
}
template <int N>
DK_HD Dual<N> operator-(double b, const Dual<N>& a) {
  Dual<N> r = -a;
  r.v += b;
  return r;
}
template <int N>
DK_HD Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] * b;
  return r;
}
template <int N>
DK_HD Dual<N> operator*(double b, const Dual<N>& a) {
  return a * b;
}
template <int N>
DK_HD Dual<N> operator/(const Dual<N>& a, double b) {
  return a * (1.0 / b);
}
template <int N>
DK_HD Dual<N> operator/(double a, const Dual<N>& b) {
  Dual<N> r;
  double ib = 1.0 / b.v;
  r.v = a * ib;
  double f = -r.v * ib;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = b.g[i] * f;
  return r;
}
template <int N>
DK_HD Dual<N> dsqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = sqrt(a.v);
  double h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] * h;
  return r;
}
template <int N>
DK_HD Dual<N> dexp(const Dual<N>& a) {
  Dual<N> r;
  r.v = exp(a.v);
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] * r.v;
  return r;
}
template <int N>
DK_HD void dsincos(const Dual<N>& a, Dual<N>* s, Dual<N>* c) {
  double sv, cv;
  sincos(a.v, &sv, &cv);
  s->v = sv;
  c->v = cv;
#pragma unroll
  for (int i = 0; i < N; i++) {
    s->g[i] = a.g[i] * cv;
    c->g[i] = -a.g[i] * sv;
  }
}
template <int N>
DK_HD double val(const Dual<N>& a) {
  return a.v;
}

DK_HD double dsqrt(double a) { return sqrt(a); }
DK_HD double dexp(double a) { return exp(a); }
DK_HD void dsincos(double a, double* s, double* c) { sincos(a, s, c); }
DK_HD double val(double a) { return a; }
template <class T>
struct Num;
template <>
struct Num<double> {
  static DK_HD double make(double v) { return v; }
};
template <int N>
struct Num<Dual<N>> {
  static DK_HD Dual<N> make(double v) { return dconst<N>(v); }
};
template <class T>
DK_HD T tconst(double v) {
  return Num<T>::make(v);
}

// Dual number in up to three directions whose partials outside the mask M are
// zero by construction, so that operations skip them: a layer matrix mixes
// P-wave quantities (alpha only) with S-wave ones (beta only), and in the
// tangent in c only k carries a partial; the compiler cannot drop a product
// with a zero partial on its own (0 * x is not 0 for every x).  The callers
// choose the directions: (alpha, beta) in the kernel pass, (alpha, beta,
// thickness) in the VJP, c alone in the tangent sweeps.
template <int M>
struct SD {
  double v;
  double g[3];
};
template <int M>
DK_HD constexpr bool sd_has(int k) {
  return ((M >> k) & 1) != 0;
}
template <int M>
DK_HD SD<M> sd_var(double v) {  // v as the variable of direction M (one bit)
  SD<M> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < 3; k++) r.g[k] = sd_has<M>(k) ? 1.0 : 0.0;
  return r;
}
template <int A, int B>
DK_HD SD<A | B> operator+(const SD<A>& a, const SD<B>& b) {
  SD<A | B> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (sd_has<A>(k) && sd_has<B>(k)) r.g[k] = a.g[k] + b.g[k];
    else if (sd_has<A>(k)) r.g[k] = a.g[k];
    else if (sd_has<B>(k)) r.g[k] = b.g[k];
    else r.g[k] = 0.0;
  }
  return r;
}
template <int A>
DK_HD SD<A> operator-(const SD<A>& a) {
  SD<A> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < 3; k++) r.g[k] = sd_has<A>(k) ? -a.g[k] : 0.0;
  return r;
}
template <int A, int B>
DK_HD SD<A | B> operator-(const SD<A>& a, const SD<B>& b) {
  SD<A | B> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (sd_has<A>(k) && sd_has<B>(k)) r.g[k] = a.g[k] - b.g[k];
    else if (sd_has<A>(k)) r.g[k] = a.g[k];
    else if (sd_has<B>(k)) r.g[k] = -b.g[k];
    else r.g[k] = 0.0;
  }
  return r;
}
template <int A, int B>
DK_HD SD<A | B> operator*(const SD<A>& a, const SD<B>& b) {
  SD<A | B> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (sd_has<A>(k) && sd_has<B>(k)) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
    else if (sd_has<A>(k)) r.g[k] = a.g[k] * b.v;
    else if (sd_has<B>(k)) r.g[k] = a.v * b.g[k];
    else r.g[k] = 0.0;
  }
  return r;
}
template <int A>
DK_HD SD<A> operator+(const SD<A>& a, double b) {
  SD<A> r = a;
  r.v += b;
  return r;
}
template <int A>
DK_HD SD<A> operator+(double b, const SD<A>& a) {
  return a + b;
}
template <int A>
DK_HD SD<A> operator-(const SD<A>& a, double b) {
  SD<A> r = a;
  r.v -= b;
  return r;
}
template <int A>
DK_HD SD<A> operator-(double b, const SD<A>& a) {
  SD<A> r = -a;
  r.v += b;
  return r;
}
template <int A>
DK_HD SD<A> operator*(const SD<A>& a, double b) {
  SD<A> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < 3; k++) r.g[k] = sd_has<A>(k) ? a.g[k] * b : 0.0;
  return r;
}
template <int A>
DK_HD SD<A> operator*(double b, const SD<A>& a) {
  return a * b;
}
template <int A>
DK_HD SD<A> operator/(double a, const SD<A>& b) {
  SD<A> r;
  double ib = 1.0 / b.v;
  r.v = a * ib;
  double f = -r.v * ib;
#pragma unroll
  for (int k = 0; k < 3; k++) r.g[k] = sd_has<A>(k) ? b.g[k] * f : 0.0;
  return r;
}
template <int A>
DK_HD double val(const SD<A>& a) {
  return a.v;
}

// ----------------------------------------------------------------- layer matrix
// Vertical wavenumber sqrt((k+kx)|k-kx|) written as sqrt(|k^2-kx^2|): the two
// forms agree in exact arithmetic and the latter keeps the derivative simple.
template <class T>
DK_HD T vwave(const T& wvno, const T& kx) {
  T s = (wvno + kx) * (wvno - kx);
  if (val(s) < 0.0) s = -s;
  return dsqrt(s);
}

// The three functions of one wave type in a layer, of X = k^2 - kx^2 (signed):
//   C = cosh(sqrt(X) d),  S = sinh(sqrt(X) d)/sqrt(X),  R = sqrt(X) sinh(sqrt(X) d),
// entire functions of X (for X < 0 they are cos, sin/., -.sin of sqrt(-X) d).
// For X > 0 the values are scaled by exp(-sqrt(X) d), which is returned as the
// exponent ex (surfdisp96.f:923-962 var).  dC, dS, dR are the derivatives of the
// entire functions with respect to X, scaled the same way: the exponent
// scaling multiplies the whole layer matrix, so for the root's implicit
// derivative it is a constant (its own derivative times F vanishes at F = 0).
// *scale receives exp(-ex) (1 for X <= 0): the factor the layer's exponent
// scaling stands for, of which a layer matrix takes the product over its two
// wave types.  (exp(-2p) is kept where the reference's var drops it for
// p >= 16, a change below 1e-13 relative.)
DK_HD double hyp3(double X, double d, double* C, double* S, double* R, double* dC, double* dS,
                  double* dR, double* scale) {
  double ex = 0.0, sc = 1.0, iX = 0.0;
  if (X > 0.0) {
    const double y = rsqrt_nr(X), ra = X * y, p = ra * d;
    sc = exp_neg(p);
    const double fac = sc * sc;
    const double sh = 0.5 - 0.5 * fac;
    *C = 0.5 + 0.5 * fac;
    *S = sh * y;
    *R = ra * sh;
    ex = p;
    iX = y * y;
  } else if (X < 0.0) {
    const double y = rsqrt_nr(-X), ra = -X * y, p = ra * d;
    double sp, cp;
    sincos_phase(p, &sp, &cp);
    *C = cp;
    *S = sp * y;
    *R = -ra * sp;
    iX = -(y * y);
  } else {
    *C = 1.0;
    *S = d;
    *R = 0.0;
  }
  *dC = 0.5 * d * (*S);
  *dR = 0.5 * ((*S) + d * (*C));
  double u = X * d * d;
  if (fabs(u) < 0.1) {
    double d3 = d * d * d;
    *dS = sc * d3 * (1.0 / 6.0 + u * (1.0 / 60.0 + u * (1.0 / 1680.0 + u * (1.0 / 90720.0))));
  } else {
    *dS = (d * (*C) - (*S)) * (0.5 * iX);
  }
  *scale = sc;
  return ex;
}

DK_HD double hyp3_t(const double& X, double d, double* S, double* R, double* ex, double* scale) {
  double C, dC, dS, dR;
  *ex = hyp3(X, d, &C, S, R, &dC, &dS, &dR, scale);
  return C;
}
// C, S, R of hyp3 as one value, for the generic layer matrix.
template <class T>
struct Hyp3 {
  T C, S, R;
};
DK_HD Hyp3<double> hyp3_v(double X, double d, double* ex, double* scale) {
  Hyp3<double> h;
  h.C = hyp3_t(X, d, &h.S, &h.R, ex, scale);
  return h;
}
template <int N, class D>
DK_HD Hyp3<Dual<N>> hyp3_v(const Dual<N>& X, const D& d, double* ex, double* scale) {
  Hyp3<Dual<N>> h;
  h.C = hyp3_t(X, d, &h.S, &h.R, ex, scale);
  return h;
}
template <int M>
DK_HD Hyp3<SD<M>> hyp3_v(const SD<M>& X, double d, double* ex, double* scale) {
  double C, Sv, Rv, dC, dS, dR;
  *ex = hyp3(X.v, d, &C, &Sv, &Rv, &dC, &dS, &dR, scale);
  Hyp3<SD<M>> h;
  h.C.v = C;
  h.S.v = Sv;
  h.R.v = Rv;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    h.C.g[k] = sd_has<M>(k) ? X.g[k] * dC : 0.0;
    h.S.g[k] = sd_has<M>(k) ? X.g[k] * dS : 0.0;
    h.R.g[k] = sd_has<M>(k) ? X.g[k] * dR : 0.0;
  }
  return h;
}
template <int M, int MD>
DK_HD Hyp3<SD<M | MD>> hyp3_v(const SD<M>& X, const SD<MD>& d, double* ex, double* scale) {
  double C, Sv, Rv, dC, dS, dR;
  *ex = hyp3(X.v, d.v, &C, &Sv, &Rv, &dC, &dS, &dR, scale);
  Hyp3<SD<M | MD>> h;
  h.C.v = C;
  h.S.v = Sv;
  h.R.v = Rv;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double xg = sd_has<M>(k) ? X.g[k] : 0.0, dg = sd_has<MD>(k) ? d.g[k] : 0.0;
    if (sd_has<M>(k) && sd_has<MD>(k)) {
      h.C.g[k] = xg * dC + dg * Rv;
      h.S.g[k] = xg * dS + dg * C;
      h.R.g[k] = xg * dR + dg * X.v * C;
    } else if (sd_has<M>(k)) {
      h.C.g[k] = xg * dC;
      h.S.g[k] = xg * dS;
      h.R.g[k] = xg * dR;
    } else if (sd_has<MD>(k)) {
      h.C.g[k] = dg * Rv;
      h.S.g[k] = dg * C;
      h.R.g[k] = dg * X.v * C;
    } else {
      h.C.g[k] = h.S.g[k] = h.R.g[k] = 0.0;
    }
  }
  return h;
}

// X and d both carry derivatives: with the exponent scaling held constant,
// dC/dd = R, dS/dd = C, dR/dd = X C.
template <int N>
DK_HD Dual<N> hyp3_t(const Dual<N>& X, const Dual<N>& d, Dual<N>* S, Dual<N>* R, double* ex,
                     double* scale) {
  double C, Sv, Rv, dC, dS, dR;
  *ex = hyp3(X.v, d.v, &C, &Sv, &Rv, &dC, &dS, &dR, scale);
  Dual<N> c;
  c.v = C;
  S->v = Sv;
  R->v = Rv;
#pragma unroll
  for (int i = 0; i < N; i++) {
    c.g[i] = X.g[i] * dC + d.g[i] * Rv;
    S->g[i] = X.g[i] * dS + d.g[i] * C;
    R->g[i] = X.g[i] * dR + d.g[i] * X.v * C;
  }
  return c;
}
template <int N>
DK_HD Dual<N> hyp3_t(const Dual<N>& X, double d, Dual<N>* S, Dual<N>* R, double* ex,
                     double* scale) {
  double C, Sv, Rv, dC, dS, dR;
  *ex = hyp3(X.v, d, &C, &Sv, &Rv, &dC, &dS, &dR, scale);
  Dual<N> c;
  c.v = C;
  S->v = Sv;
  R->v = Rv;
#pragma unroll
  for (int i = 0; i < N; i++) {
    c.g[i] = X.g[i] * dC;
    S->g[i] = X.g[i] * dS;
    R->g[i] = X.g[i] * dR;
  }
  return c;
}

// One layer: Dunkin's reduced compound matrix ca (surfdisp96.f:808-831 plus
// var and dnka), scaled by exp(-exa). a,b,rho are the flattened layer values,
// d its flattened thickness (a double, or a Dual<N> to differentiate in d).
// visit(j, i, value) receives every entry ca[j][i] once as it is formed (so a
// consumer can reduce the matrix without storing it); layer_ca stores them.
template <class T, class D = double, class V, class TA = T, class TB = T, class TR = T>
DK_HD void layer_ca_visit(const T& wvno, const T& wvno2, const TA& a, const TB& b,
                          const TR& rho, const D& d, double omega, V&& visit) {
  const auto xka = omega / a;
  const auto xkb = omega / b;
  const auto t = b * (1.0 / omega);
  const auto gammk = 2.0 * t * t;
  const auto gam = gammk * wvno2;
  // ---- var (surfdisp96.f:868-985): cosp=C_P, w=S_P, x=R_P; cosq=C_S, y=S_S, z=R_S
  double pex, sex, sP, sS;
  const auto hP = hyp3_v((wvno + xka) * (wvno - xka), d, &pex, &sP);
  const auto hS = hyp3_v((wvno + xkb) * (wvno - xkb), d, &sex, &sS);
  const auto &cosp = hP.C, &w = hP.S, &x = hP.R;
  const auto &cosq = hS.C, &y = hS.S, &z = hS.R;
  // exp(-(pex + sex)); the reference's var drops it for pex + sex >= 60 (< 1e-26)
  const double a0 = sP * sS;
  const auto cpcq = cosp * cosq;
  const auto cpy = cosp * y;
  const auto cpz = cosp * z;
  const auto cqw = cosq * w;
  const auto cqx = cosq * x;
  const auto xy = x * y;
  const auto xz = x * z;
  const auto wy = w * y;
  const auto wz = w * z;
  // ---- dnka (surfdisp96.f:1018-1062)
  const auto gamm1 = gam - 1.0;
  const auto twgm1 = gam + gamm1;
  const auto gmgmk = gam * gammk;
  const auto gmgm1 = gam * gamm1;
  const auto gm1sq = gamm1 * gamm1;
  const auto irho = 1.0 / rho;
  const auto rho2 = rho * rho;
  const auto irho2 = irho * irho;
  const auto a0pq = a0 - cpcq;
  const auto tt = -2.0 * wvno2;
  {
    const auto c00 = cpcq - 2.0 * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
    visit(0, 0, c00);
    visit(4, 4, c00);
    visit(2, 2, a0 + 2.0 * (cpcq - c00));
  }
  {
    const auto c01 = (wvno2 * cpy - cqx) * irho;
    visit(0, 1, c01);
    visit(3, 4, c01);
  }
  {
    const auto c02 = -(twgm1 * a0pq + gammk * xz + wvno2 * gamm1 * wy) * irho;
    visit(0, 2, c02);
    visit(2, 4, tt * c02);
  }
  {
    const auto c03 = (cpz - wvno2 * cqw) * irho;
    visit(0, 3, c03);
    visit(1, 4, c03);
  }
  visit(0, 4, -(2.0 * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) * irho2);
  {
    const auto c10 = (gmgmk * cpz - gm1sq * cqw) * rho;
    visit(1, 0, c10);
    visit(4, 3, c10);
  }
  visit(1, 1, cpcq);
  visit(3, 3, cpcq);
  {
    const auto c12 = gammk * cpz - gamm1 * cqw;
    visit(1, 2, c12);
    visit(2, 3, tt * c12);
  }
  visit(1, 3, -wz);
  {
    const auto c30 = (gm1sq * cpy - gmgmk * cqx) * rho;
    visit(3, 0, c30);
    visit(4, 1, c30);
  }
  visit(3, 1, -xy);
  {
    const auto c32 = gamm1 * cpy - gammk * cqx;
    visit(3, 2, c32);
    visit(2, 1, tt * c32);
  }
  visit(4, 0, -(2.0 * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2);
  {
    const auto c42 =
        -(gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * xz + gamm1 * gm1sq * wy) * rho;
    visit(4, 2, c42);
    visit(2, 0, tt * c42);
  }
}

// Every entry of the layer matrix is rho^p times a rho-free factor, p =
// RHO_POW[j][i] (dnka's rho, 1/rho, rho^2, 1/rho^2 factors), so its rho
// derivative is p * ca[j][i] / rho; likewise the half-space vector with powers
// HALF_RHO_POW.
__host__ __device__ constexpr int rho_pow(int j, int i) {
  constexpr int P[5][5] = {{0, -1, -1, -1, -2},
                           {1, 0, 0, 0, -1},
                           {1, 0, 0, 0, -1},
                           {1, 0, 0, 0, -1},
                           {2, 1, 1, 1, 0}};
  return P[j][i];
}
__host__ __device__ constexpr int half_rho_pow(int i) { return i == 0 ? 2 : (i == 4 ? 0 : 1); }

template <class T, class D = double>
DK_HD void layer_ca(const T& wvno, const T& wvno2, const T& a, const T& b, const T& rho,
                    const D& d, double omega, T ca[5][5]) {
  layer_ca_visit<T, D>(wvno, wvno2, a, b, rho, d, omega,
                       [&](int j, int i, const T& v) { ca[j][i] = v; });
}

// Half-space 5-vector (surfdisp96.f:777-802).
template <class T>
DK_HD void halfspace_e(const T& wvno, const T& wvno2, const T& a, const T& b, const T& rho,
                       double omega, T e[5]) {
  T xka = omega / a;
  T xkb = omega / b;
  T ra = vwave(wvno, xka);
  T rb = vwave(wvno, xkb);
  T t = b * (1.0 / omega);
  T gammk = 2.0 * t * t;
  T gam = gammk * wvno2;
  T gamm1 = gam - 1.0;
  T rarb = ra * rb;
  e[0] = rho * rho * (gamm1 * gamm1 - gam * gammk * rarb);
  e[1] = -(rho * ra);
  e[2] = rho * (gamm1 - gammk * rarb);
  e[3] = rho * rb;
  e[4] = wvno2 - rarb;
}

// ----------------------------------------------------------------- value-only recursion
// Exact power-of-two normalisation: scale e by 2^-k with k = exponent of max|e|.
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

// Layers are given as double4 (d, a, b, rho) of the flattened model; index
// L-1 is the half-space. Returns the mantissa; *ex receives the exponent.
template <class LayerFn>
DK_HD double secular(double c, double omega, int L, LayerFn lay, int* ex) {
  if (omega < 1.0e-4) omega = 1.0e-4;
  double wvno = omega / c;
  double wvno2 = wvno * wvno;
  double e[5];
  double4 h = lay(L - 1);
  halfspace_e<double>(wvno, wvno2, (double)h.y, (double)h.z, (double)h.w, omega, e);
  int E = pow2_normalize(e);
  for (int m = L - 2; m >= 0; m--) {
    double4 l = lay(m);
    double ca[5][5];
    layer_ca<double>(wvno, wvno2, (double)l.y, (double)l.z, (double)l.w, (double)l.x, omega, ca);
    double ee[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double cr = 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) cr = fma(e[j], ca[j][i], cr);
      ee[i] = cr;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) e[i] = ee[i];
    E += pow2_normalize(e);
  }
  *ex = E;
  return e[0];
}

}  // namespace dk
