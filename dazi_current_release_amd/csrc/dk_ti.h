// Transversely isotropic (TI) Rayleigh-wave depth kernels of one (column, period),
// evaluated at a given phase velocity, for the dc/dA, dc/dL splice that feeds the
// Gc/Gs rows (reference: code/src_inv_iso_joint/depthkernelTI.f90:93-162 with the
// CPS engine tregn96.f).
//
// The quantity reproduced is tregn96's: eigenfunctions at the *given* (omega, k)
// (the float32 phase velocity of the isotropic root search), built from
//   * a compound-matrix sweep of the half-space radiation constraints up to every
//     interface (tregn96.f:1834-1990 up),
//   * a propagator sweep of the traction-free surface solution (1,0,0,0) down to
//     every interface (tregn96.f:3561-3709 down),
//   * their symplectic contraction at each interface, normalised by the surface
//     minor (tregn96.f:1559-1832 svfunc),
// followed by the closed-form layer energy integrals (tregn96.f:3777-4245 energy,
// intijr), the group-velocity normalisation, the sphericity back-mapping
// (sprayl, tregn96.f:1235-1299) and the float32 chain of depthkernelTI.
//
// Layer algebra (TI eigenvalues and eigenvector components x11..x42, NP, NSV)
// follows tregn96.f:3166-3361 (gettiegn) and the eigenvector matrix E and its
// inverse G of evalg (:2989-3164).  The layer propagator H = H_P + H_S is written
// per wave type as [u v~] M [a~ b~]^T with det M = 1, so its second compound is
//   e^{-(p+q)} (u^v~)(a~^b~)^T / NP^2 + same for S + mixed(H_P, H_S),
// evaluated without cancellation; this is the matrix tregn96 expands term by
// term in dnka_tregn (:1992-2987), reduced to 5x5 by the same minor identity.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "dk_fastmath.h"

namespace dk {
namespace ti {

// ----------------------------------------------------------------- complex double
struct cx {
  double r, i;
};
DK_HD cx mk(double r, double i = 0.0) { return cx{r, i}; }
DK_HD cx operator+(cx a, cx b) { return cx{a.r + b.r, a.i + b.i}; }
DK_HD cx operator-(cx a, cx b) { return cx{a.r - b.r, a.i - b.i}; }
DK_HD cx operator-(cx a) { return cx{-a.r, -a.i}; }
DK_HD cx operator*(cx a, cx b) { return cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
DK_HD cx operator*(double s, cx a) { return cx{s * a.r, s * a.i}; }
DK_HD cx operator*(cx a, double s) { return cx{s * a.r, s * a.i}; }
DK_HD cx operator/(cx a, cx b) {
  // scaled division (Smith's algorithm)
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    return cx{(a.r + a.i * t) / d, (a.i - a.r * t) / d};
  }
  double t = b.r / b.i, d = b.r * t + b.i;
  return cx{(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
DK_HD cx operator/(cx a, double s) { return cx{a.r / s, a.i / s}; }
DK_HD cx operator/(double s, cx b) { return cx{s, 0.0} / b; }
DK_HD double cabs_(cx a) { return hypot(a.r, a.i); }
// a / b through 1/|b|^2: one division, for divisors of physical size (vertical
// wavenumbers and their sums), where |b|^2 neither overflows nor underflows
DK_HD cx cdiv_fast(cx a, cx b) {
  const double inv = 1.0 / (b.r * b.r + b.i * b.i);
  return cx{(a.r * b.r + a.i * b.i) * inv, (a.i * b.r - a.r * b.i) * inv};
}
// Scalar-generic helpers: the layer algebra below is written once for real
// (double) and complex (cx) eigenvector components.
DK_HD cx mk(cx a) { return a; }
DK_HD double re(double a) { return a; }
DK_HD double re(cx a) { return a.r; }
DK_HD double mad(double a, double b, double c) { return fma(a, b, c); }
DK_HD cx mad(cx a, cx b, cx c) { return a * b + c; }
template <class T>
DK_HD T lift(double v);
template <>
DK_HD double lift<double>(double v) { return v; }
template <>
DK_HD cx lift<cx>(double v) { return cx{v, 0.0}; }
DK_HD cx csqrt_(cx z) {
  // principal square root; a zero imaginary part takes its sign (branch cut)
  if (z.r == 0.0 && z.i == 0.0) return cx{0.0, z.i};
  double m = hypot(z.r, z.i);
  double t = sqrt(0.5 * (m + fabs(z.r)));
  if (z.r >= 0.0) return cx{t, z.i / (2.0 * t)};
  return cx{fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}

// ----------------------------------------------------------------- layer model
struct Layer {  // flattened TI layer (float32 values of the reference, as double)
  double d, A, C, F, L, N, rho;
  // period-independent constants: 1/C, 1/L, eta = F/(A-2L), F/C,
  // sqrt(A/rho), sqrt(C/rho), sqrt(L/rho) and F/eta
  double iC, iL, eta, FiC, ah, av, bv, Feta;
};

// The constants of a layer whose d, A, C, F, L, N, rho are set.
DK_HD void layer_constants(Layer* m) {
  m->iC = 1.0 / m->C;
  m->iL = 1.0 / m->L;
  m->eta = m->F / (m->A - 2.0 * m->L);
  m->FiC = m->F / m->C;
  m->ah = sqrt(m->A / m->rho);
  m->av = sqrt(m->C / m->rho);
  m->bv = sqrt(m->L / m->rho);
  m->Feta = m->F / m->eta;
}

// Coefficients of gettiegn's characteristic equation L^2 - bb L + cc = 0 in
// L = nu^2 (tregn96.f:3212-3244), with the eigenvector algebra's a..f.
struct Quad {
  double a, b, c, d, e, f, bb, cc;
};

DK_HD Quad quad(const Layer& m, double omega, double wvno) {
  const double om2 = omega * omega, wvno2 = wvno * wvno;
  const double iC = m.iC, iL = m.iL;
  Quad q;
  q.a = wvno * m.F * iC;
  q.b = iC;
  q.c = -m.rho * om2 + wvno2 * (m.A - m.F * m.F * iC);
  q.d = -wvno;
  q.e = iL;
  q.f = -m.rho * om2;
  const double ddef = wvno2 - m.rho * om2 * iL;
  const double aabc = wvno2 * m.A * iC - m.rho * om2 * iC;
  q.bb = 2.0 * (q.a * q.d) + q.e * q.c + q.f * q.b;
  q.cc = ddef * aabc;
  return q;
}

// Eigenvector components from the two roots (gettiegn, tregn96.f:3301-3340),
// each column normalised so that its leading component is k.  T is double or cx.
template <class T>
DK_HD void eigvec(const Quad& q, double wvno, T L1, T L2, T x[8]) {
  const double bcaa = q.b * q.c + q.a * q.a, ddef2 = q.d * q.d + q.e * q.f;
  T x12 = lift<T>(q.b * q.d - q.a * q.e);
  T x22 = q.b * L2 - lift<T>(q.e * bcaa);
  T x32 = L2 - lift<T>(q.a * q.d + q.c * q.e);
  T x42 = -(q.a * L2) + lift<T>(q.d * bcaa);
  T x11 = -(q.e * L1) + lift<T>(q.b * ddef2);
  T x21 = lift<T>(q.b * q.d - q.a * q.e);
  T x31 = q.d * L1 - lift<T>(q.a * ddef2);
  T x41 = -(L1 - lift<T>(q.a * q.d) - lift<T>(q.b * q.f));
  if (wvno != 0.0) {
    T z = wvno / x11;
    x11 = x11 * z;
    x21 = x21 * z;
    x31 = x31 * z;
    x41 = x41 * z;
    z = wvno / x22;
    x12 = x12 * z;
    x22 = x22 * z;
    x32 = x32 * z;
    x42 = x42 * z;
  }
  x[0] = x11;
  x[1] = x21;
  x[2] = x31;
  x[3] = x41;
  x[4] = x12;
  x[5] = x22;
  x[6] = x32;
  x[7] = x42;
}

// Eigen-solution of one layer (gettiegn) at real (omega, k).  The squared
// vertical wavenumbers L1 (P) and L2 (SV) are the roots of a real quadratic.
// When its discriminant is non-negative, both roots are real and every
// quantity of the up and down sweeps is real: nu = sqrt(L) is real
// (evanescent) or i*sqrt(-L) (oscillatory).  ok = 0 flags a negative
// discriminant: L1, L2 are then a complex-conjugate pair (near-isotropic
// layers where the P and SV wavenumbers nearly coincide), which EigC handles.
struct Hyp;
struct HypC;

struct Eig {
  using T = double;   // eigenvector component type
  using H = Hyp;      // its layer phase functions
  double L1, L2;      // nu_P^2, nu_SV^2
  double nP, nS;      // |nu_P|, |nu_SV|
  double rnP, rnS;    // 1/|nu_P|, 1/|nu_SV|
  int iP, iS;         // nu imaginary
  double x11, x21, x31, x41, x12, x22, x32, x42;
  double NP, NSV;
  double rNP, rNSV;   // 1/NP, 1/NSV
  int ok;
};

// (wvno > 0, iwv2 = 1/wvno^2)
DK_HD void eigen(const Layer& m, double omega, double wvno, double iwv2, Eig* o) {
  const double wvno2 = wvno * wvno;
  const Quad q = quad(m, omega, wvno);
  const double bb = q.bb, cc = q.cc;
  const double disc = bb * bb - 4.0 * cc;
  o->ok = disc >= 0.0;
  const double s = sqrt(fmax(disc, 0.0));
  double L1 = (bb + s) * 0.5, L2;
  L2 = (L1 != 0.0) ? cc / L1 : (bb - s) * 0.5;
  if (fabs(wvno2 - L2) < fabs(wvno2 - L1)) {
    double t = L1;
    L1 = L2;
    L2 = t;
  }
  o->L1 = L1;
  o->L2 = L2;
  o->iP = L1 < 0.0;
  o->iS = L2 < 0.0;
  const double aP = fabs(L1), aS = fabs(L2);
  o->rnP = rsqrt_nr(aP);
  o->rnS = rsqrt_nr(aS);
  o->nP = aP > 0.0 ? aP * o->rnP : 0.0;
  o->nS = aS > 0.0 ? aS * o->rnS : 0.0;
  // eigvec's columns before their normalisation z = k/x11 (k/x22), and
  // N = z^2 Nr of the normalised ones: one division per wave gives both z
  // and 1/N (R = 1/(x11 Nr): z = k Nr R, 1/N = x11^3 R / k^2)
  const double bcaa = q.b * q.c + q.a * q.a, ddef2 = q.d * q.d + q.e * q.f;
  const double x12 = q.b * q.d - q.a * q.e;
  const double x22 = q.b * L2 - q.e * bcaa;
  const double x32 = L2 - (q.a * q.d + q.c * q.e);
  const double x42 = -(q.a * L2) + q.d * bcaa;
  const double x11 = -(q.e * L1) + q.b * ddef2;
  const double x21 = q.b * q.d - q.a * q.e;
  const double x31 = q.d * L1 - q.a * ddef2;
  const double x41 = -(L1 - q.a * q.d - q.b * q.f);
  const double NPr = x11 * x41 - x21 * x31, NSr = x12 * x42 - x22 * x32;
  const double R1 = 1.0 / (x11 * NPr), R2 = 1.0 / (x22 * NSr);
  const double z1 = wvno * NPr * R1, z2 = wvno * NSr * R2;
  o->x11 = x11 * z1;
  o->x21 = x21 * z1;
  o->x31 = x31 * z1;
  o->x41 = x41 * z1;
  o->x12 = x12 * z2;
  o->x22 = x22 * z2;
  o->x32 = x32 * z2;
  o->x42 = x42 * z2;
  o->NP = o->x11 * o->x41 - o->x21 * o->x31;
  o->NSV = o->x12 * o->x42 - o->x22 * o->x32;
  o->rNP = x11 * x11 * (x11 * R1) * iwv2;
  o->rNSV = x22 * x22 * (x22 * R2) * iwv2;
}

// Complex vertical wavenumbers of the eigen-solution.
DK_HD cx nu_P(const Eig& g) { return g.iP ? cx{0.0, g.nP} : cx{g.nP, 0.0}; }
DK_HD cx nu_S(const Eig& g) { return g.iS ? cx{0.0, g.nS} : cx{g.nS, 0.0}; }
// 1/(2 nu N): nu real -> real, nu = i|nu| -> -i/(2|nu|N)
DK_HD cx inv2nuN_P(const Eig& g) {
  const double v = (0.5 * g.rNP) * g.rnP;
  return g.iP ? cx{0.0, -v} : cx{v, 0.0};
}
DK_HD cx inv2nuN_S(const Eig& g) {
  const double v = (0.5 * g.rNSV) * g.rnS;
  return g.iS ? cx{0.0, -v} : cx{v, 0.0};
}
// 1/(2 nu)
DK_HD cx inv2nu_P(const Eig& g) { return g.iP ? cx{0.0, -0.5 * g.rnP} : cx{0.5 * g.rnP, 0.0}; }
DK_HD cx inv2nu_S(const Eig& g) { return g.iS ? cx{0.0, -0.5 * g.rnS} : cx{0.5 * g.rnS, 0.0}; }

// General complex eigen-solution: the form tregn96 evaluates in complex*16
// throughout.  Layers with real roots take exactly Eig's values (same
// branches); a complex-conjugate pair takes L1 = (bb + i sqrt(-disc))/2,
// L2 = cc/L1, the P/SV labels by the smaller |k^2 - L|, and the principal
// square roots (Re nu > 0) (tregn96.f:3246-3300).
struct EigC {
  using T = cx;
  using H = HypC;
  cx rp, rs;  // nu_P, nu_SV
  cx x11, x21, x31, x41, x12, x22, x32, x42;
  cx NP, NSV;
  cx rNP, rNSV;  // 1/NP, 1/NSV
  int ok;     // always 1
};

DK_HD void eigen(const Layer& m, double omega, double wvno, double iwv2, EigC* o) {
  Eig g;
  eigen(m, omega, wvno, iwv2, &g);
  o->ok = 1;
  if (g.ok) {
    o->rp = nu_P(g);
    o->rs = nu_S(g);
    o->x11 = mk(g.x11);
    o->x21 = mk(g.x21);
    o->x31 = mk(g.x31);
    o->x41 = mk(g.x41);
    o->x12 = mk(g.x12);
    o->x22 = mk(g.x22);
    o->x32 = mk(g.x32);
    o->x42 = mk(g.x42);
    o->NP = mk(g.NP);
    o->NSV = mk(g.NSV);
    o->rNP = mk(g.rNP);
    o->rNSV = mk(g.rNSV);
    return;
  }
  const double wvno2 = wvno * wvno;
  const Quad q = quad(m, omega, wvno);
  const double disc = q.bb * q.bb - 4.0 * q.cc;
  cx L1 = cx{0.5 * q.bb, 0.5 * sqrt(fmax(-disc, 0.0))};
  cx L2 = q.cc / L1;
  if (cabs_(mk(wvno2) - L2) < cabs_(mk(wvno2) - L1)) {
    cx t = L1;
    L1 = L2;
    L2 = t;
  }
  o->rp = csqrt_(L1);
  o->rs = csqrt_(L2);
  if (o->rp.r < 0.0) o->rp = -o->rp;
  if (o->rs.r < 0.0) o->rs = -o->rs;
  cx x[8];
  eigvec(q, wvno, L1, L2, x);
  o->x11 = x[0];
  o->x21 = x[1];
  o->x31 = x[2];
  o->x41 = x[3];
  o->x12 = x[4];
  o->x22 = x[5];
  o->x32 = x[6];
  o->x42 = x[7];
  o->NP = x[0] * x[3] - x[1] * x[2];
  o->NSV = x[4] * x[7] - x[5] * x[6];
  o->rNP = cdiv_fast(mk(1.0), o->NP);
  o->rNSV = cdiv_fast(mk(1.0), o->NSV);
}

DK_HD cx nu_P(const EigC& g) { return g.rp; }
DK_HD cx nu_S(const EigC& g) { return g.rs; }
DK_HD cx inv2nuN_P(const EigC& g) { return cdiv_fast(mk(0.5), g.rp * g.NP); }
DK_HD cx inv2nuN_S(const EigC& g) { return cdiv_fast(mk(0.5), g.rs * g.NSV); }
DK_HD cx inv2nu_P(const EigC& g) { return cdiv_fast(mk(0.5), g.rp); }
DK_HD cx inv2nu_S(const EigC& g) { return cdiv_fast(mk(0.5), g.rs); }

// Eigenvector matrix E (columns P+, S+, P-, S-) and its inverse G (evalg).
template <class G_>
DK_HD void eig_matrices(const G_& g, cx E[4][4], cx G[4][4]) {
  const cx rp = nu_P(g), rs = nu_S(g);
  const cx iP = inv2nuN_P(g), iS = inv2nuN_S(g);
  const auto hP = 0.5 * g.rNP, hS = 0.5 * g.rNSV;  // nu/(2 nu N)
  G[0][0] = mk(g.x41 * hP);
  G[1][0] = g.x42 * iS;
  G[2][0] = mk(g.x41 * hP);
  G[3][0] = -(g.x42 * iS);
  G[0][1] = -(g.x31 * iP);
  G[1][1] = mk(-(g.x32 * hS));
  G[2][1] = g.x31 * iP;
  G[3][1] = mk(-(g.x32 * hS));
  G[0][2] = mk(-(g.x21 * hP));
  G[1][2] = -(g.x22 * iS);
  G[2][2] = mk(-(g.x21 * hP));
  G[3][2] = g.x22 * iS;
  G[0][3] = g.x11 * iP;
  G[1][3] = mk(g.x12 * hS);
  G[2][3] = -(g.x11 * iP);
  G[3][3] = mk(g.x12 * hS);
  E[0][0] = mk(g.x11);
  E[1][0] = g.x21 * rp;
  E[2][0] = mk(g.x31);
  E[3][0] = g.x41 * rp;
  E[0][1] = g.x12 * rs;
  E[1][1] = mk(g.x22);
  E[2][1] = g.x32 * rs;
  E[3][1] = mk(g.x42);
  E[0][2] = mk(g.x11);
  E[1][2] = -(g.x21 * rp);
  E[2][2] = mk(g.x31);
  E[3][2] = -(g.x41 * rp);
  E[0][3] = -(g.x12 * rs);
  E[1][3] = mk(g.x22);
  E[2][3] = -(g.x32 * rs);
  E[3][3] = mk(g.x42);
}

// Exponentially scaled cosh/sinh of the P and SV phases (varsv), real:
// nu real: cosh, nu sinh, sinh/nu of p = nu d scaled by e^{-p}; nu = i|nu|:
// cos, -|nu| sin, sin/|nu| of |nu| d.  Per wave also its scale e^{-Re p}
// (scl) and the exponential e^{-p} itself (e), which the layer integrals use.
struct Hyp {
  double cosp, rsinp, sinpr, cossv, rsinsv, sinsvr;
  double pex, svex;
  double sclp, sclsv;
  cx eP, eS;
};

DK_HD void hyp_wave(int imag, double n, double rn, double dm, double* co, double* rsi,
                    double* sir, double* ex, double* scl, cx* e) {
  const double p = n * dm;
  if (!imag) {
    const double s = exp_neg(p);
    const double fac = (p < 15.0) ? s * s : 0.0;  // var drops e^{-2p} from p = 15
    const double sh = 0.5 - 0.5 * fac;
    *co = 0.5 + 0.5 * fac;
    *rsi = n * sh;
    *sir = (p < 1.0e-5 && n < 1.0e-5) ? dm : sh * rn;
    *ex = p;
    *scl = s;
    *e = cx{s, 0.0};
  } else {
    double s, c;
    sincos_phase(p, &s, &c);
    *co = c;
    *rsi = -n * s;
    *sir = (n < 1.0e-5) ? dm : s * rn;
    *ex = 0.0;
    *scl = 1.0;
    *e = cx{c, -s};
  }
}

DK_HD void hyper(const Eig& g, double dm, Hyp* h) {
  hyp_wave(g.iP, g.nP, g.rnP, dm, &h->cosp, &h->rsinp, &h->sinpr, &h->pex, &h->sclp, &h->eP);
  hyp_wave(g.iS, g.nS, g.rnS, dm, &h->cossv, &h->rsinsv, &h->sinsvr, &h->svex, &h->sclsv,
           &h->eS);
}

// The same for complex nu (varsv's general form, tregn96.f:3363-3440): with
// p = nu d, cosh p, nu sinh p and sinh p / nu scaled by e^{-Re p}.
struct HypC {
  cx cosp, rsinp, sinpr, cossv, rsinsv, sinsvr;
  double pex, svex;
  double sclp, sclsv;
  cx eP, eS;
};

// (Re nu >= 0, so Re p >= 0)
DK_HD void hyp_wave(cx n, double dm, cx* co, cx* rsi, cx* sir, double* ex, double* scl, cx* e) {
  const cx p = n * dm;
  double s, c;
  sincos_phase(fabs(p.i), &s, &c);
  if (p.i < 0.0) s = -s;
  const cx ep = cx{0.5 * c, 0.5 * s}, em = cx{0.5 * c, -0.5 * s};
  const double sc = exp_neg(p.r);
  const double fac = (p.r < 15.0) ? sc * sc : 0.0;
  const cx sh = ep - fac * em;
  *co = ep + fac * em;
  *rsi = n * sh;
  *sir = (fabs(p.r) < 1.0e-5 && cabs_(n) < 1.0e-5) ? mk(dm) : cdiv_fast(sh, n);
  *ex = p.r;
  *scl = sc;
  *e = cx{sc * c, -(sc * s)};
}

DK_HD void hyper(const EigC& g, double dm, HypC* h) {
  hyp_wave(g.rp, dm, &h->cosp, &h->rsinp, &h->sinpr, &h->pex, &h->sclp, &h->eP);
  hyp_wave(g.rs, dm, &h->cossv, &h->rsinsv, &h->sinsvr, &h->svex, &h->sclsv, &h->eS);
}

// Per wave type, H = [u v~] M [a~ b~]^T / N with
//   P: u=(x11,0,x31,0) v~=(0,x21,0,x41) a~=(x41,0,-x21,0) b~=(0,-x31,0,x11),
//      M = [[cosp, sinpr],[rsinp, cosp]]
//   S: u=(x12,0,x32,0) v~=(0,x22,0,x42) a~=(x42,0,-x22,0) b~=(0,-x32,0,x12),
//      M = [[cossv, rsinsv],[sinsvr, cossv]]
// u and v~ have disjoint supports (rows 0,2 and 1,3), as have a~ and b~
// (columns 0,2 and 1,3), so H(i,j) = w_i * M(i%2, j%2) * z_j / N with
// w = (x11, x21, x31, x41) and z = (x41, -x31, -x21, x11) for P.
template <class T>
struct Wave {
  T w[4], z[4];
  T m00, m01, m10;  // divided by N (m11 = m00)
};

template <class G_, class H_, class T = typename G_::T>
DK_HD void waves(const G_& g, const H_& h, double sP, double sS, Wave<T>* P, Wave<T>* S) {
  P->w[0] = g.x11;
  P->w[1] = g.x21;
  P->w[2] = g.x31;
  P->w[3] = g.x41;
  P->z[0] = g.x41;
  P->z[1] = -g.x31;
  P->z[2] = -g.x21;
  P->z[3] = g.x11;
  S->w[0] = g.x12;
  S->w[1] = g.x22;
  S->w[2] = g.x32;
  S->w[3] = g.x42;
  S->z[0] = g.x42;
  S->z[1] = -g.x32;
  S->z[2] = -g.x22;
  S->z[3] = g.x12;
  const T iNP = sP * g.rNP, iNS = sS * g.rNSV;
  P->m00 = h.cosp * iNP;
  P->m01 = h.sinpr * iNP;
  P->m10 = h.rsinp * iNP;
  S->m00 = h.cossv * iNS;
  S->m01 = h.rsinsv * iNS;
  S->m10 = h.sinsvr * iNS;
}

template <class T>
DK_HD T wave_entry(const Wave<T>& w, int i, int j) {
  const T m = ((i & 1) == 0) ? (((j & 1) == 0) ? w.m00 : w.m01)
                             : (((j & 1) == 0) ? w.m10 : w.m00);
  return w.w[i] * m * w.z[j];
}

// out = H v for the real layer propagator H (hska: the real part of the
// complex one), P part scaled by sP and S part by sS.  Per wave,
// (H v)_i = w_i (M (z . v))_i%2 with the even and odd halves of z . v.
template <class G_, class H_>
DK_HD void propagate(const G_& g, const H_& h, double sP, double sS, const double v[4],
                     double out[4]) {
  using T = typename G_::T;
  Wave<T> P, S;
  waves(g, h, sP, sS, &P, &S);
  auto wave = [&](const Wave<T>& w, T o[4]) {
    const T te = w.z[0] * v[0] + w.z[2] * v[2], to = w.z[1] * v[1] + w.z[3] * v[3];
    const T ev = w.m00 * te + w.m01 * to, od = w.m10 * te + w.m00 * to;
    o[0] = w.w[0] * ev;
    o[1] = w.w[1] * od;
    o[2] = w.w[2] * ev;
    o[3] = w.w[3] * od;
  };
  T oP[4], oS[4];
  wave(P, oP);
  wave(S, oS);
#pragma unroll
  for (int i = 0; i < 4; i++) out[i] = re(oP[i] + oS[i]);
}

// Minor index pairs in the order (12),(13),(14),(23),(24),(34); the reduced
// 5-vector drops (23) = -(14).
__host__ __device__ constexpr int PI0(int k) { return k < 3 ? 0 : (k < 5 ? 1 : 2); }
__host__ __device__ constexpr int PI1(int k) { return k == 0 ? 1 : k == 1 ? 2 : k == 2 ? 3 : k == 3 ? 2 : 3; }
// reduced index -> pair index
__host__ __device__ constexpr int RED(int r) { return r < 3 ? r : r + 1; }

// e <- e * CA with CA the reduced 5x5 compound of the layer propagator scaled
// by e^{-(pex+svex)} (dnka_tregn): pure-wave compounds
// dfac (u^v~)(a~^b~)^T / N^2 plus the mixed terms of H_P and H_S; row (14)
// absorbs (23) = -(14).  CA is never formed.  With the vector's coefficients
// as the antisymmetric 4x4 C (C01 = e0, C02 = e1, C03 = e2, C12 = -e2,
// C13 = e3, C23 = e4), the mixed part of column pair (k,l) is
// sum_{i<j} C_ij (HP_ik HS_jl - HP_il HS_jk + HS_ik HP_jl - HS_il HP_jk)
// = A_kl - A_lk with A = HP^T C HS, and H_ik = w_i M(i%2, k%2) z_k per wave
// gives A_kl = zP_k zS_l G[k%2][l%2], G from the four parity sums of
// wP_i C_ij wS_j.  The pure-wave part is (sum_r c_r (u^v~)_r) (a~^b~)_s per
// wave, where (u^v~) is (w0 w1, 0, w0 w3, -w1 w2, 0, w2 w3) over the pairs
// (01, 02, 03, 12, 13, 23), and the same for (a~^b~) with z.
template <class G_, class H_, class T = typename G_::T>
DK_HD void compound_step(const G_& g, const H_& h, T e[5]) {
  const double ex = h.pex + h.svex;
  const double dfac = (ex > 35.0) ? 0.0 : h.sclp * h.sclsv;  // e^{-ex}
  Wave<T> P, S;
  waves(g, h, 1.0, 1.0, &P, &S);
  const T e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3], e4 = e[4];
  const T* wP = P.w;
  const T* wS = S.w;
  // pure-wave part: c . (u^v~) per wave, times dfac / N^2
  const T iNP2 = dfac * (g.rNP * g.rNP), iNS2 = dfac * (g.rNSV * g.rNSV);
  auto cw = [&](const T* w) { return e0 * (w[0] * w[1]) + e2 * (w[0] * w[3] + w[1] * w[2]) +
                                     e4 * (w[2] * w[3]); };
  const T CP = cw(wP) * iNP2, CS = cw(wS) * iNS2;
  // mixed part: parity sums E[a][b] of wP_i C_ij wS_j (i%2 = a, j%2 = b)
  const T E00 = e1 * (wP[0] * wS[2] - wP[2] * wS[0]);
  const T E11 = e3 * (wP[1] * wS[3] - wP[3] * wS[1]);
  const T E01 = wP[0] * (e0 * wS[1] + e2 * wS[3]) + wP[2] * (e2 * wS[1] + e4 * wS[3]);
  const T E10 = -(wP[1] * (e0 * wS[0] + e2 * wS[2]) + wP[3] * (e2 * wS[0] + e4 * wS[2]));
  // M(a, b) of each wave: (0,0) = (1,1) = m00, (0,1) = m01, (1,0) = m10
  auto mP = [&](int a, int b) { return a == b ? P.m00 : (a == 0 ? P.m01 : P.m10); };
  auto mS = [&](int a, int b) { return a == b ? S.m00 : (a == 0 ? S.m01 : S.m10); };
  T F[2][2], Gp[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int be = 0; be < 2; be++)
      F[a][be] = (a == 0 ? E00 : E10) * mS(0, be) + (a == 0 ? E01 : E11) * mS(1, be);
#pragma unroll
  for (int al = 0; al < 2; al++)
#pragma unroll
    for (int be = 0; be < 2; be++) Gp[al][be] = mP(0, al) * F[0][be] + mP(1, al) * F[1][be];
  auto A = [&](int k, int l) { return (P.z[k] * S.z[l]) * Gp[k & 1][l & 1]; };
  const T* zP = P.z;
  const T* zS = S.z;
  e[0] = CP * (zP[0] * zP[1]) + CS * (zS[0] * zS[1]) + (A(0, 1) - A(1, 0));
  e[1] = A(0, 2) - A(2, 0);
  e[2] = CP * (zP[0] * zP[3]) + CS * (zS[0] * zS[3]) + (A(0, 3) - A(3, 0));
  e[3] = A(1, 3) - A(3, 1);
  e[4] = CP * (zP[2] * zP[3]) + CS * (zS[2] * zS[3]) + (A(2, 3) - A(3, 2));
}

// Half-space radiation constraint: minors of rows P+ and S+ of G (evalg gbr),
// real parts as up takes them (tregn96.f:1926-1930).
template <class G_, class T = typename G_::T>
DK_HD void halfspace_constraint(const G_& g, T cd[5]) {
  cx E[4][4], G[4][4];
  eig_matrices(g, E, G);
  cx CG[6];
  for (int r = 0; r < 6; r++) {
    int i = PI0(r), j = PI1(r);
    CG[r] = G[0][i] * G[1][j] - G[0][j] * G[1][i];
  }
  cd[0] = lift<T>(CG[0].r);
  cd[1] = lift<T>(CG[1].r);
  cd[2] = lift<T>(CG[2].r);
  cd[3] = lift<T>(CG[4].r);
  cd[4] = lift<T>(CG[5].r);
}

// ffunc, gfunc, h1func, h2func (tregn96.f:1385-1473), with ea = e^{-nua dm}
// and eb = e^{-nub dm} given: every exponential they take is a product of
// these (e^{-2 nub dm} = eb^2, e^{-(nua+nub) dm} = ea eb), with the
// reference's cut-offs on Re of the exponent.
DK_HD double cabs2_(cx a) { return a.r * a.r + a.i * a.i; }
// (i2nub = 1/(2 nub))
DK_HD cx ffunc(cx nub, double dm, cx eb, cx i2nub) {
  if (cabs2_(nub) < 1.0e-16) return mk(dm);
  cx ex = (nub.r * dm < 40.0) ? eb * eb : mk(0.0);
  return (mk(1.0) - ex) * i2nub;
}
DK_HD cx gfunc(cx nub, double dm, cx eb) {
  return (nub.r * dm < 75.0) ? eb * dm : mk(0.0);
}
DK_HD cx h1func(cx nua, cx nub, double dm, cx ea, cx eb) {
  const cx s = nub + nua;
  if (cabs2_(s) < 1.0e-16) return mk(dm);
  cx ex = (s.r * dm < 40.0) ? ea * eb : mk(0.0);
  return cdiv_fast(mk(1.0) - ex, s);
}
DK_HD cx h2func(cx nua, cx nub, double dm, cx ea, cx eb) {
  if (cabs2_(nub - nua) < 1.0e-16) return mk(dm);
  cx exp_ = (nua.r * dm < 40.0) ? ea : mk(0.0);
  cx exq = (nub.r * dm < 40.0) ? eb : mk(0.0);
  return cdiv_fast(exq - exp_, nua - nub);
}

// The six integrals out[t] = Re(a_i a_j A + b_i b_j B + (a_i b_j + b_i a_j) C)
// of intijr's (i, j) = (0,0) (0,2) (1,1) (1,3) (2,2) (3,3), with (A, B, C) =
// (Ae, Be, Ce) for even rows and (Ao, Bo, Co) for odd rows.
DK_HD void integrals_out(const EigC& g, cx Ae, cx Be, cx Ce, cx Ao, cx Bo, cx Co,
                         double out[6]) {
  const cx ra = nu_P(g), rb = nu_S(g);
  const cx a[4] = {mk(g.x11), g.x21 * ra, mk(g.x31), g.x41 * ra};
  const cx b[4] = {g.x12 * rb, mk(g.x22), g.x32 * rb, mk(g.x42)};
  const int I[6] = {0, 0, 1, 1, 2, 3}, J[6] = {0, 2, 1, 3, 2, 3};
#pragma unroll
  for (int t = 0; t < 6; t++) {
    const int i = I[t], j = J[t];
    const bool ev = (i & 1) == 0;
    const cx v = a[i] * a[j] * (ev ? Ae : Ao) + b[i] * b[j] * (ev ? Be : Bo) +
                 (a[i] * b[j] + b[i] * a[j]) * (ev ? Ce : Co);
    out[t] = v.r;
  }
}
// Real eigenvector components: a = (x11, x21 nuP, x31, x41 nuP) and
// b = (x12 nuS, x22, x32 nuS, x42), so each product above is real up to one
// factor: b_i b_j carries nuS^2 and the cross term nuS on even rows, a_i a_j
// nuP^2 and the cross term nuP on odd rows (nu^2 real).
DK_HD void integrals_out(const Eig& g, cx Ae, cx Be, cx Ce, cx Ao, cx Bo, cx Co, double out[6]) {
  const cx ra = nu_P(g), rb = nu_S(g);
  const double nuP2 = (g.iP ? -g.nP : g.nP) * g.nP, nuS2 = (g.iS ? -g.nS : g.nS) * g.nS;
  const double Ae_ = Ae.r, Be_ = nuS2 * Be.r, Ce_ = (rb * Ce).r;
  const double Ao_ = nuP2 * Ao.r, Bo_ = Bo.r, Co_ = (ra * Co).r;
  auto ev = [&](double ai, double aj, double bi, double bj) {
    return ai * aj * Ae_ + bi * bj * Be_ + (ai * bj + bi * aj) * Ce_;
  };
  auto od = [&](double ai, double aj, double bi, double bj) {
    return ai * aj * Ao_ + bi * bj * Bo_ + (ai * bj + bi * aj) * Co_;
  };
  out[0] = ev(g.x11, g.x11, g.x12, g.x12);
  out[1] = ev(g.x11, g.x31, g.x12, g.x32);
  out[2] = od(g.x21, g.x21, g.x22, g.x22);
  out[3] = od(g.x21, g.x41, g.x22, g.x42);
  out[4] = ev(g.x31, g.x31, g.x32, g.x32);
  out[5] = od(g.x41, g.x41, g.x42, g.x42);
}

// Layer integrals INT11, INT13, INT22, INT24, INT33, INT44 of the field whose
// up-going amplitudes come from the bottom interface value yb and down-going
// ones from the top value yt (intijr); half = 1 for the half-space, else eP,
// eS are e^{-nu dm} of the P and SV waves (unused for the half-space).
// With the eigenvector matrix E written as E[i] = (a_i, b_i, s_i a_i, -s_i b_i),
// a = (x11, x21 nuP, x31, x41 nuP), b = (x12 nuS, x22, x32 nuS, x42), s_i = +1
// for even rows and -1 for odd rows, and every integral pairing rows of equal
// parity, the sum over the ten wave-pair products of intijr collapses to
//   Re(a_i a_j A_s + b_i b_j B_s + (a_i b_j + b_i a_j) C_s).
template <class G_>
DK_HD void layer_integrals(const G_& g, double dm, const double yt[4], const double yb[4],
                           int half, cx eP, cx eS, double out[6]) {
  const cx ra = nu_P(g), rb = nu_S(g);
  // rows of G = E^-1: 1/(2 nu N) and nu/(2 nu N) = 1/(2N)
  const cx iP = inv2nuN_P(g), iS = inv2nuN_S(g);
  const auto hP = 0.5 * g.rNP, hS = 0.5 * g.rNSV;
  const cx i2a = inv2nu_P(g), i2b = inv2nu_S(g);
  // amplitudes: up-going P, SV from yb; down-going P, SV from yt
  const cx k0 = mk(hP * (g.x41 * yb[0] - g.x21 * yb[2])) + iP * (g.x11 * yb[3] - g.x31 * yb[1]);
  const cx k1 = iS * (g.x42 * yb[0] - g.x22 * yb[2]) + mk(hS * (g.x12 * yb[3] - g.x32 * yb[1]));
  const cx k2 = mk(hP * (g.x41 * yt[0] - g.x21 * yt[2])) + iP * (g.x31 * yt[1] - g.x11 * yt[3]);
  const cx k3 = iS * (g.x22 * yt[2] - g.x42 * yt[0]) + mk(hS * (g.x12 * yt[3] - g.x32 * yt[1]));
  cx Ae, Be, Ce, Ao, Bo, Co;
  if (half) {
    Ae = Ao = (k2 * k2) * i2a;
    Be = Bo = (k3 * k3) * i2b;
    Ce = Co = -cdiv_fast(k2 * k3, ra + rb);
  } else {
    const cx FA = ffunc(ra, dm, eP, i2a), GA = gfunc(ra, dm, eP), FB = ffunc(rb, dm, eS, i2b),
             GB = gfunc(rb, dm, eS);
    const cx H1 = h1func(ra, rb, dm, eP, eS), H2 = h2func(ra, rb, dm, eP, eS);
    const cx p00 = (k0 * k0 + k2 * k2) * FA, p02 = 2.0 * (GA * (k0 * k2));
    const cx p11 = (k1 * k1 + k3 * k3) * FB, p13 = 2.0 * (GB * (k1 * k3));
    const cx p01 = H1 * (k0 * k1 - k2 * k3), p21 = H2 * (k2 * k1 - k0 * k3);
    Ae = p00 + p02;
    Ao = p00 - p02;
    Be = p11 - p13;
    Bo = p11 + p13;
    Ce = p01 + p21;
    Co = p01 - p21;
  }
  integrals_out(g, Ae, Be, Ce, Ao, Bo, Co, out);
}

// Exact power-of-two normalisations; they return the natural log of the
// factor divided out, as cnormc / rnormc do for their max-norm factor.
DK_HD double norm5(double e[5]) {
  double t = 0.0;
  for (int i = 0; i < 5; i++) t = fmax(t, fabs(e[i]));
  if (!(t >= 1.0e-40) || !isfinite(t)) return 0.0;
  int k;
  frexp(t, &k);
  for (int i = 0; i < 5; i++) e[i] = ldexp(e[i], -k);
  return (double)k * 0.69314718055994530942;
}
DK_HD double norm5(cx e[5]) {
  double t = 0.0;
  for (int i = 0; i < 5; i++) t = fmax(t, cabs_(e[i]));
  if (!(t >= 1.0e-40) || !isfinite(t)) return 0.0;
  int k;
  frexp(t, &k);
  for (int i = 0; i < 5; i++) e[i] = cx{ldexp(e[i].r, -k), ldexp(e[i].i, -k)};
  return (double)k * 0.69314718055994530942;
}
DK_HD double norm4(double e[4]) {
  double t = 0.0;
  for (int i = 0; i < 4; i++) t = fmax(t, fabs(e[i]));
  if (!(t >= 1.0e-40) || !isfinite(t)) return 0.0;
  int k;
  frexp(t, &k);
  for (int i = 0; i < 4; i++) e[i] = ldexp(e[i], -k);
  return (double)k * 0.69314718055994530942;
}

// ----------------------------------------------------------------- model prep
// Earth flattening of the TI model (sphere_tdisp96, radius 6371: A, C, L, N
// scale by tmp^-0.275, rho by tmp^-2.275, F is left unchanged) and the
// sphericity factors vtp of bldsph (radius 6370 on the flattened thicknesses).
// Inputs are the refined layers (thickness, TA, TL, TF, rho; TC = TA, TN = TL),
// half-space last; float32 storage as in the reference.
DK_HD void flatten(int L, const float* thk, const float* TA, const float* TL, const float* TF,
                   const float* rho, Layer* out, float* vtp) {
#pragma clang fp contract(off)
  const double ar = 6371.0;
  const double e_r = (double)(-2.275f), e_m = (double)(-0.2750f);
  double r0 = ar;
  for (int i = 0; i < L; i++) {
    float td = (i == L - 1) ? 1.0f : thk[i];
    double r1 = r0 - (double)td;
    double z0 = ar * log(ar / r0);
    double z1 = ar * log(ar / r1);
    float tdf = (float)(z1 - z0);
    double tmp = (ar + ar) / (r0 + r1);
    float rf = (float)((double)rho[i] * pow(tmp, e_r));
    double fm = pow(tmp, e_m);
    float af = (float)((double)TA[i] * fm);
    float lf = (float)((double)TL[i] * fm);
    out[i].d = (i == L - 1) ? 0.0 : (double)tdf;
    out[i].A = af;
    out[i].C = af;
    out[i].F = TF[i];
    out[i].L = lf;
    out[i].N = lf;
    out[i].rho = rf;
    layer_constants(&out[i]);
    r0 = r1;
  }
  const double ar2 = 6370.0;
  r0 = ar2;
  for (int i = 0; i < L; i++) {
    double zd = (i == L - 1) ? 1.0 : out[i].d;
    double r1 = r0 * exp(-zd / ar2);
    double tmp = (ar2 + ar2) / (r0 + r1);
    vtp[i] = (float)tmp;
    r0 = r1;
  }
}

// ----------------------------------------------------------------- one (column, period)
// Scratch per layer and thread (element (m, q) at scratch[(m * N + q) * stride]):
// the up-sweep compound vector (5 values of T, real and imaginary parts
// adjacent for cx) and its accumulated log scale (1), and the unnormalised
// kernel factors facah, facbv, facn (3).  The interface eigenfunction lives in
// registers (each layer needs its top and bottom value).
template <class T>
struct Slots {
  static constexpr int W = sizeof(T) / sizeof(double);
  static constexpr int CD = 0, EX = 5 * W, F = 5 * W + 1, N = 5 * W + 4;
};
enum { Q_N = Slots<double>::N, Q_NC = Slots<cx>::N };
// Scratch accesses: on the device they are non-temporal (streamed past the
// caches, which keep the layer records instead).
DK_HD void scr_st(double* p, double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
DK_HD double scr_ld(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
// Kernel factor q (facah, facbv, facn) of layer m in the scratch of a solve
// in T arithmetic.
template <class T>
DK_HD double factor(const double* scratch, long stride, int m, int q) {
  return scr_ld(scratch + (long)(m * Slots<T>::N + Slots<T>::F + q) * stride);
}
// chksiz's float rounding of a kernel value
DK_HD float kernel_f32(double x) { return fabs(x) < 1.0e-36 ? 0.0f : (float)x; }

// solve<Eig> runs in real arithmetic and returns -2 when a layer's eigenvalues
// form a complex-conjugate pair; solve<EigC> takes every layer in complex
// arithmetic (the same values where the roots are real).  On return 0 the
// scratch holds each layer's kernel factors (factor<T>), and tregn96's kernels
// after the sphericity back-mapping are, top layer first,
//   dcdah = f32(facah vtp isc), dcdbv = f32(facbv vtp isc), dcdn = f32(facn inrm)
// with f32 = kernel_f32 and vtp the layer's float sphericity factor.
// head(m) returns the flattened Layer with d, A, C, F, L, N, rho, iC and iL
// set, tail(m, &layer) sets the rest of its constants.  Returns 0, -1 when the
// surface minor vanishes, or -2 as above.
template <class G_, class HeadFn, class TailFn>
DK_HD int solve(int L, HeadFn head, TailFn tail, double omega, double c, double* scratch,
                long stride, double* isc_out, double* inrm_out) {
  using T = typename G_::T;
  using Q = Slots<T>;
  auto P = [&](int m, int q) -> double* { return scratch + (long)(m * Q::N + q) * stride; };
  auto S = [&](int m, int q) -> double { return scr_ld(P(m, q)); };
  auto put_cd = [&](int m, const T* e) {
    for (int q = 0; q < 5; q++) {
      if constexpr (Q::W == 1) {
        scr_st(P(m, Q::CD + q), re(e[q]));
      } else {
        scr_st(P(m, Q::CD + 2 * q), re(e[q]));
        scr_st(P(m, Q::CD + 2 * q + 1), e[q].i);
      }
    }
  };
  // real part of compound entry q at interface m (svfunc reads real parts)
  auto cd_re = [&](int m, int q) -> double { return S(m, Q::CD + Q::W * q); };
  const double wvno = omega / c, wvno2 = wvno * wvno, om2 = omega * omega;
  const double iwv = 1.0 / wvno, iwv2 = iwv * iwv;
  G_ g;
  typename G_::H h;
  int ok = 1;
  // ---- up: compound constraint vectors
  T cd[5];
  eigen(head(L - 1), omega, wvno, iwv2, &g);
  ok &= g.ok;
  halfspace_constraint(g, cd);
  put_cd(L - 1, cd);
  scr_st(P(L - 1, Q::EX), 0.0);
  double exsum = 0.0;
  // each layer is loaded one iteration ahead, so that its latency overlaps the
  // layer before it
  Layer ahead = head(L >= 2 ? L - 2 : 0);
  for (int m = L - 2; m >= 0; m--) {
    const Layer ly = ahead;
    ahead = head(m >= 1 ? m - 1 : 0);
    // the previous layer's vector is stored after the next layer's reads are
    // issued (a read issued after stores waits for them too); the vector and
    // its scale are this iteration's inputs, so they are in registers anyway.
    // At m = L - 2 this rewrites the half-space entries with their own values.
    put_cd(m + 1, cd);
    scr_st(P(m + 1, Q::EX), exsum);
    eigen(ly, omega, wvno, iwv2, &g);
    ok &= g.ok;
    hyper(g, ly.d, &h);
    compound_step(g, h, cd);
    const double exn = norm5(cd);
    exsum = exsum + h.pex + h.svex + exn;
  }
  if (L >= 2) {
    put_cd(0, cd);
    scr_st(P(0, Q::EX), exsum);
  }
  if (!ok) return -2;
  // ---- surface eigenfunction: ur = Re(cd(1,3)/cd(1,2)), f1213 = -Re cd(1,2)
  const double exe1 = S(0, Q::EX);
  const double f1213 = -cd_re(0, 1);
  if (!(f1213 != 0.0) || !isfinite(f1213)) return -1;
  const double if1213 = 1.0 / f1213;
  double ur0;
  if constexpr (Q::W == 1) {
    ur0 = S(0, Q::CD + 2) / S(0, Q::CD + 1);
  } else {
    ur0 = (cx{S(0, Q::CD + 4), S(0, Q::CD + 5)} / cx{S(0, Q::CD + 2), S(0, Q::CD + 3)}).r;
  }
  double yt[4] = {ur0, 1.0, 0.0, 0.0};
  // Q dispersion of c (gammap: Qp = 150, Qs = 50, fref = 1 Hz), accumulated
  // over the layers and normalised at the end; used only by the sphericity
  // factor tm below
  const double pi = 3.141592653589493;
  const double qai = (double)(1.0f / 150.0f), qbi = (double)(1.0f / 50.0f);
  const double lw = log(omega / (2.0 * pi * 1.0)) / pi;
  double dcq = 0.0;
  // ---- down sweep, interface eigenfunctions and layer integrals
  double vv[4] = {1.0, 0.0, 0.0, 0.0};
  double exd = 0.0;
  double sumi0 = 0, sumi1 = 0, sumi2 = 0, sumi3 = 0;
  ahead = head(0);
  for (int m = 0; m < L; m++) {
    Layer ly = ahead;
    const int n = m + 1 < L ? m + 1 : L - 1;
    ahead = head(n);
    // this layer's remaining constants (used last) and the up sweep's values at
    // interface m+1 are read before the propagator work, which hides their
    // latency
    tail(m, &ly);
    const double cd1 = cd_re(n, 0), cd2 = cd_re(n, 1), cd3 = cd_re(n, 2), cd5 = cd_re(n, 3),
                 cd6 = cd_re(n, 4), exn_up = S(n, Q::EX);
    eigen(ly, omega, wvno, iwv2, &g);
    double yb[4] = {0.0, 0.0, 0.0, 0.0};  // interface m+1 (zero below the half-space top)
    if (m < L - 1) {
      hyper(g, ly.d, &h);
      double sP = 1.0, sS = 1.0, cpex;
      if (h.pex > h.svex) {
        sS = (h.pex - h.svex > 40.0) ? 0.0 : exp_neg(h.pex - h.svex);
        cpex = h.pex;
      } else {
        sP = (h.svex - h.pex > 40.0) ? 0.0 : exp_neg(h.svex - h.pex);
        cpex = h.svex;
      }
      double a0[4];
      propagate(g, h, sP, sS, vv, a0);
      double ex2 = norm4(a0);
      exd = exd + cpex + ex2;
      for (int i = 0; i < 4; i++) vv[i] = a0[i];
      // contraction at interface m+1 (svfunc)
      const double cd4 = -cd3;
      double tz1 = -vv[3], tz2 = -vv[2], tz3 = vv[1], tz4 = vv[0];
      double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
      double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
      double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
      double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
      double ext = exd + exn_up - exe1;
      if (ext > -80.0 && ext < 80.0) {
        const double fact = exp_neg(-ext) * if1213;
        yb[0] = uu1 * fact;
        yb[1] = uu2 * fact;
        yb[2] = uu3 * fact;
        yb[3] = uu4 * fact;
      }
    }
    double I[6];
    layer_integrals(g, ly.d, yt, yb, m == L - 1, h.eP, h.eS, I);
    const double INT11 = I[0], INT13 = I[1], INT22 = I[2], INT24 = I[3], INT33 = I[4],
                 INT44 = I[5];
    // energy (getmat coefficients)
    const double rho = ly.rho;
    const double ah = ly.ah, av = ly.av, bv = ly.bv;
    const double eta = ly.eta;
    const double a12 = -wvno, a14 = ly.iL, a21 = wvno * ly.FiC, a23 = ly.iC;
    const double URUR = INT11, UZUZ = INT22;
    const double DURDUR = a12 * a12 * INT22 + 2.0 * a12 * a14 * INT24 + a14 * a14 * INT44;
    const double DUZDUZ = a21 * a21 * INT11 + 2.0 * a21 * a23 * INT13 + a23 * a23 * INT33;
    const double URDUZ = a21 * INT11 + a23 * INT13;
    const double UZDUR = a12 * INT22 + a14 * INT24;
    sumi0 = sumi0 + rho * (URUR + UZUZ);
    sumi1 = sumi1 + ly.L * UZUZ + ly.A * URUR;
    sumi2 = sumi2 + ly.L * UZDUR - ly.F * URDUZ;
    sumi3 = sumi3 + ly.L * DURDUR + ly.C * DUZDUZ;
    const double facah = rho * ah * (URUR - 2.0 * eta * URDUZ * iwv);
    const double facav = rho * av * DUZDUZ * iwv2;
    const double facbv = rho * bv * (UZUZ + 2.0 * UZDUR * iwv + DURDUR * iwv2 + 4.0 * eta * URDUZ * iwv);
    const double facn = -ly.Feta * URDUZ * iwv;
    scr_st(P(m, Q::F + 0), facah);
    scr_st(P(m, Q::F + 1), facbv);
    scr_st(P(m, Q::F + 2), facn);
    dcq = dcq + lw * (facbv * bv * qbi) + lw * (facav * av * qai + facah * ah * qai);
#pragma unroll
    for (int i = 0; i < 4; i++) yt[i] = yb[i];
  }
  const double ugr = (wvno * sumi1 + sumi2) / (omega * sumi0);
  const double nrm = ugr * sumi0;
  const double dc = dcq / nrm;
  const double cq = omega / wvno + dc;
  const double ar = 6370.0;
  const double tm = sqrt(1.0 + (cq / (2.0 * ar * omega)) * (cq / (2.0 * ar * omega)));
  const double tm3 = tm * tm * tm;
  (void)om2;
  *inrm_out = 1.0 / nrm;
  *isc_out = *inrm_out / tm3;
  return 0;
}

// depthkernelTI's chain to the moduli in float32 (depthkernelTI.f90:152-159):
// dc/dA = dc/dalpha_h / (2 rho vp) - F/(A-2L)^2 dc/deta and
// dc/dL = dc/dbeta_v / (2 rho vs) + 2F/(A-2L)^2 dc/deta of one refined layer,
// from tregn96's float32 kernels dah, dbv, dn.
DK_HD void chain_AL(float rrho, float rvp, float rvs, float TA, float TL, float TF, float dah,
                    float dbv, float dn, float* dA, float* dL) {
#pragma clang fp contract(off)
  float den = TA - 2.0f * TL;
  float den2 = den * den;
  *dA = 0.5f / (rrho * rvp) * dah - TF / den2 * dn;
  *dL = 0.5f / (rrho * rvs) * dbv + 2.0f * TF / den2 * dn;
}

// depthkernelTI splice in float32 (depthkernelTI.f90:152-162): interval j sums
// its nsub[j] sublayers, each as acc = (acc + dc/dA*TA) + dc/dL*TL (the
// reference's left-to-right order).
DK_HD float splice_add(float acc, float TA, float TL, float dA, float dL) {
#pragma clang fp contract(off)
  return (acc + dA * TA) + dL * TL;
}

}  // namespace ti
}  // namespace dk
