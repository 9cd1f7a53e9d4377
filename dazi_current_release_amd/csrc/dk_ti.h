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

#ifndef DK_HD
#define DK_HD __host__ __device__ inline
#endif

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
DK_HD double cabs_(cx a) { return hypot(a.r, a.i); }
DK_HD cx csqrt_(cx z) {
  // principal square root; a zero imaginary part takes its sign (branch cut)
  if (z.r == 0.0 && z.i == 0.0) return cx{0.0, z.i};
  double m = hypot(z.r, z.i);
  double t = sqrt(0.5 * (m + fabs(z.r)));
  if (z.r >= 0.0) return cx{t, z.i / (2.0 * t)};
  return cx{fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}
DK_HD cx cexp_(cx z) {
  double e = exp(z.r), s, c;
  sincos(z.i, &s, &c);
  return cx{e * c, e * s};
}

// ----------------------------------------------------------------- layer model
struct Layer {  // flattened TI layer (float32 values of the reference, as double)
  double d, A, C, F, L, N, rho;
};

// Eigen-solution of one layer (gettiegn).
struct Eig {
  cx rp, rsv;
  cx x11, x21, x31, x41, x12, x22, x32, x42;
  cx NP, NSV;
};

DK_HD void eigen(const Layer& m, double omega, double wvno, Eig* o) {
  const double om2 = omega * omega, wvno2 = wvno * wvno;
  cx a = mk(wvno * m.F / m.C);
  cx b = mk(1.0 / m.C);
  cx c = mk(-m.rho * omega * omega + wvno * wvno * (m.A - m.F * m.F / m.C));
  cx d = mk(-wvno);
  cx e = mk(1.0 / m.L);
  cx f = mk(-m.rho * omega * omega);
  cx ddef = mk(wvno * wvno - m.rho * omega * omega / m.L);
  cx aabc = mk(wvno * wvno * m.A / m.C - m.rho * omega * omega / m.C);
  cx bb = 2.0 * (a * d) + e * c + f * b;
  cx cc = ddef * aabc;
  cx s = csqrt_(bb * bb - 4.0 * cc);
  if (s.i < 0.0) s = -s;
  cx L1, L2;
  if (bb.r < 0.0 && s.r < 0.0) {
    L2 = (bb - s) / 2.0;
    L1 = (cabs_(L2) > 0.0) ? cc / L2 : (bb + s) / 2.0;
  } else {
    L1 = (bb + s) / 2.0;
    L2 = (cabs_(L1) > 0.0) ? cc / L1 : (bb - s) / 2.0;
  }
  cx xka2 = mk(wvno2) - L1, xkb2 = mk(wvno2) - L2;
  if (cabs_(xkb2) < cabs_(xka2)) {
    cx t = L1;
    L1 = L2;
    L2 = t;
  }
  o->rp = csqrt_(L1);
  o->rsv = csqrt_(L2);
  if (o->rp.r < 0.0) o->rp = -o->rp;
  if (o->rsv.r < 0.0) o->rsv = -o->rsv;
  cx bcaa = b * c + a * a, ddef2 = d * d + e * f;
  o->x12 = b * d - a * e;
  o->x22 = b * L2 - e * bcaa;
  o->x32 = L2 - (a * d + c * e);
  o->x42 = -(a * L2) + d * bcaa;
  o->x11 = -(e * L1) + b * ddef2;
  o->x21 = b * d - a * e;
  o->x31 = d * L1 - a * ddef2;
  o->x41 = -(L1 - a * d - b * f);
  if (wvno != 0.0) {
    cx z = mk(wvno) / o->x11;
    o->x11 = o->x11 * z;
    o->x21 = o->x21 * z;
    o->x31 = o->x31 * z;
    o->x41 = o->x41 * z;
    z = mk(wvno) / o->x22;
    o->x12 = o->x12 * z;
    o->x22 = o->x22 * z;
    o->x32 = o->x32 * z;
    o->x42 = o->x42 * z;
  }
  o->NP = o->x11 * o->x41 - o->x21 * o->x31;
  o->NSV = o->x12 * o->x42 - o->x22 * o->x32;
  (void)om2;
}

// Eigenvector matrix E (columns P+, S+, P-, S-) and its inverse G (evalg).
DK_HD void eig_matrices(const Eig& g, cx E[4][4], cx G[4][4]) {
  const cx rp = g.rp, rs = g.rsv;
  cx dP = 2.0 * (rp * g.NP), dS = 2.0 * (rs * g.NSV);
  G[0][0] = g.x41 * rp / dP;
  G[1][0] = g.x42 / dS;
  G[2][0] = g.x41 * rp / dP;
  G[3][0] = -(g.x42 / dS);
  G[0][1] = -(g.x31 / dP);
  G[1][1] = -(g.x32 * rs / dS);
  G[2][1] = g.x31 / dP;
  G[3][1] = -(g.x32 * rs / dS);
  G[0][2] = -(g.x21 * rp / dP);
  G[1][2] = -(g.x22 / dS);
  G[2][2] = -(g.x21 * rp / dP);
  G[3][2] = g.x22 / dS;
  G[0][3] = g.x11 / dP;
  G[1][3] = g.x12 * rs / dS;
  G[2][3] = -(g.x11 / dP);
  G[3][3] = g.x12 * rs / dS;
  E[0][0] = g.x11;
  E[1][0] = g.x21 * rp;
  E[2][0] = g.x31;
  E[3][0] = g.x41 * rp;
  E[0][1] = g.x12 * rs;
  E[1][1] = g.x22;
  E[2][1] = g.x32 * rs;
  E[3][1] = g.x42;
  E[0][2] = g.x11;
  E[1][2] = -(g.x21 * rp);
  E[2][2] = g.x31;
  E[3][2] = -(g.x41 * rp);
  E[0][3] = -(g.x12 * rs);
  E[1][3] = g.x22;
  E[2][3] = -(g.x32 * rs);
  E[3][3] = g.x42;
}

// Exponentially scaled cosh/sinh of the P and SV phases (varsv).
struct Hyp {
  cx cosp, rsinp, sinpr, cossv, rsinsv, sinsvr;
  double pex, svex;
};

DK_HD void hyper(const Eig& g, double dm, Hyp* h) {
  cx p = g.rp * dm, q = g.rsv * dm;
  double s, c;
  h->pex = p.r;
  sincos(p.i, &s, &c);
  cx epp = cx{c / 2.0, s / 2.0}, epm = cx{c / 2.0, -s / 2.0};
  double pfac = (p.r < 15.0) ? exp(-2.0 * p.r) : 0.0;
  h->cosp = epp + pfac * epm;
  cx sinp = epp - pfac * epm;
  h->rsinp = g.rp * sinp;
  h->sinpr = (fabs(p.r) < 1.0e-5 && cabs_(g.rp) < 1.0e-5) ? mk(dm) : sinp / g.rp;
  h->svex = q.r;
  sincos(q.i, &s, &c);
  cx eqp = cx{c / 2.0, s / 2.0}, eqm = cx{c / 2.0, -s / 2.0};
  double sfac = (q.r < 15.0) ? exp(-2.0 * q.r) : 0.0;
  h->cossv = eqp + sfac * eqm;
  cx sinq = eqp - sfac * eqm;
  h->rsinsv = g.rsv * sinq;
  h->sinsvr = (fabs(q.r) < 1.0e-5 && cabs_(g.rsv) < 1.0e-5) ? mk(dm) : sinq / g.rsv;
}

// Per wave type, H = [u v~] M [a~ b~]^T / N with
//   P: u=(x11,0,x31,0) v~=(0,x21,0,x41) a~=(x41,0,-x21,0) b~=(0,-x31,0,x11),
//      M = [[cosp, sinpr],[rsinp, cosp]]
//   S: u=(x12,0,x32,0) v~=(0,x22,0,x42) a~=(x42,0,-x22,0) b~=(0,-x32,0,x12),
//      M = [[cossv, rsinsv],[sinsvr, cossv]]
struct Wave {
  cx u[4], v[4], a[4], b[4];
  cx m00, m01, m10, m11;  // already divided by N
};

DK_HD void waves(const Eig& g, const Hyp& h, double sP, double sS, Wave* P, Wave* S) {
  cx z = mk(0.0);
  cx uP[4] = {g.x11, z, g.x31, z}, vP[4] = {z, g.x21, z, g.x41};
  cx aP[4] = {g.x41, z, -g.x21, z}, bP[4] = {z, -g.x31, z, g.x11};
  cx uS[4] = {g.x12, z, g.x32, z}, vS[4] = {z, g.x22, z, g.x42};
  cx aS[4] = {g.x42, z, -g.x22, z}, bS[4] = {z, -g.x32, z, g.x12};
  for (int i = 0; i < 4; i++) {
    P->u[i] = uP[i];
    P->v[i] = vP[i];
    P->a[i] = aP[i];
    P->b[i] = bP[i];
    S->u[i] = uS[i];
    S->v[i] = vS[i];
    S->a[i] = aS[i];
    S->b[i] = bS[i];
  }
  P->m00 = sP * h.cosp / g.NP;
  P->m01 = sP * h.sinpr / g.NP;
  P->m10 = sP * h.rsinp / g.NP;
  P->m11 = P->m00;
  S->m00 = sS * h.cossv / g.NSV;
  S->m01 = sS * h.rsinsv / g.NSV;
  S->m10 = sS * h.sinsvr / g.NSV;
  S->m11 = S->m00;
}

DK_HD cx wave_entry(const Wave& w, int i, int j) {
  return w.u[i] * (w.m00 * w.a[j] + w.m01 * w.b[j]) + w.v[i] * (w.m10 * w.a[j] + w.m11 * w.b[j]);
}

// Real layer propagator (hska), P part scaled by sP and S part by sS.
DK_HD void propagator(const Eig& g, const Hyp& h, double sP, double sS, double H[4][4]) {
  Wave P, S;
  waves(g, h, sP, sS, &P, &S);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) H[i][j] = (wave_entry(P, i, j) + wave_entry(S, i, j)).r;
}

// Minor index pairs in the order (12),(13),(14),(23),(24),(34); the reduced
// 5-vector drops (23) = -(14).
__host__ __device__ constexpr int PI0(int k) { return k < 3 ? 0 : (k < 5 ? 1 : 2); }
__host__ __device__ constexpr int PI1(int k) { return k == 0 ? 1 : k == 1 ? 2 : k == 2 ? 3 : k == 3 ? 2 : 3; }
// reduced index -> pair index
__host__ __device__ constexpr int RED(int r) { return r < 3 ? r : r + 1; }

// Reduced 5x5 compound of the layer propagator, scaled by e^{-(pex+svex)} (dnka_tregn).
DK_HD void compound(const Eig& g, const Hyp& h, cx CA[5][5]) {
  double ex = h.pex + h.svex;
  double dfac = (ex > 35.0) ? 0.0 : exp(-ex);
  Wave P, S;
  waves(g, h, 1.0, 1.0, &P, &S);
  cx HP[4][4], HS[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      HP[i][j] = wave_entry(P, i, j);
      HS[i][j] = wave_entry(S, i, j);
    }
  // pure-wave compounds (u^v~)(a~^b~)/N^2, unscaled, times dfac
  cx iNP2 = mk(1.0) / (g.NP * g.NP), iNS2 = mk(1.0) / (g.NSV * g.NSV);
  cx T6[6][6];
  for (int r = 0; r < 6; r++) {
    int i = PI0(r), j = PI1(r);
    cx uvP = P.u[i] * P.v[j] - P.u[j] * P.v[i];
    cx uvS = S.u[i] * S.v[j] - S.u[j] * S.v[i];
    for (int s = 0; s < 6; s++) {
      int k = PI0(s), l = PI1(s);
      cx abP = P.a[k] * P.b[l] - P.a[l] * P.b[k];
      cx abS = S.a[k] * S.b[l] - S.a[l] * S.b[k];
      cx pure = dfac * (uvP * abP * iNP2 + uvS * abS * iNS2);
      cx mixed = HP[i][k] * HS[j][l] - HP[i][l] * HS[j][k] + HS[i][k] * HP[j][l] -
                 HS[i][l] * HP[j][k];
      T6[r][s] = pure + mixed;
    }
  }
  for (int r = 0; r < 5; r++)
    for (int s = 0; s < 5; s++) {
      cx v = T6[RED(r)][RED(s)];
      if (r == 2) v = v - T6[3][RED(s)];  // row (14) absorbs (23) = -(14)
      CA[r][s] = v;
    }
}

// Half-space radiation constraint: minors of rows P+ and S+ of G (evalg gbr).
DK_HD void halfspace_constraint(const Eig& g, cx cd[5]) {
  cx E[4][4], G[4][4];
  eig_matrices(g, E, G);
  cx CG[6];
  for (int r = 0; r < 6; r++) {
    int i = PI0(r), j = PI1(r);
    CG[r] = G[0][i] * G[1][j] - G[0][j] * G[1][i];
  }
  cd[0] = mk(CG[0].r);
  cd[1] = mk(CG[1].r);
  cd[2] = mk(CG[2].r);
  cd[3] = mk(CG[4].r);
  cd[4] = mk(CG[5].r);
}

// ffunc, gfunc, h1func, h2func (tregn96.f:1385-1473)
DK_HD cx ffunc(cx nub, double dm) {
  if (cabs_(nub) < 1.0e-08) return mk(dm);
  cx arg = nub * dm;
  cx ex = (arg.r < 40.0) ? cexp_(-2.0 * arg) : mk(0.0);
  return (mk(1.0) - ex) / (2.0 * nub);
}
DK_HD cx gfunc(cx nub, double dm) {
  cx arg = nub * dm;
  return (arg.r < 75.0) ? cexp_(-arg) * dm : mk(0.0);
}
DK_HD cx h1func(cx nua, cx nub, double dm) {
  if (cabs_(nub + nua) < 1.0e-08) return mk(dm);
  cx arg = (nua + nub) * dm;
  cx ex = (arg.r < 40.0) ? cexp_(-arg) : mk(0.0);
  return (mk(1.0) - ex) / (nub + nua);
}
DK_HD cx h2func(cx nua, cx nub, double dm) {
  if (cabs_(nub - nua) < 1.0e-08) return mk(dm);
  cx ap = nua * dm, aq = nub * dm;
  cx exp_ = (ap.r < 40.0) ? cexp_(-ap) : mk(0.0);
  cx exq = (aq.r < 40.0) ? cexp_(-aq) : mk(0.0);
  return (exq - exp_) / (nua - nub);
}

// Layer integrals INT11, INT13, INT22, INT24, INT33, INT44 of the field whose
// up-going amplitudes come from the bottom interface value yb and down-going
// ones from the top value yt (intijr); half = 1 for the half-space.
DK_HD void layer_integrals(const Eig& g, double dm, const double yt[4], const double yb[4],
                           int half, double out[6]) {
  cx E[4][4], G[4][4];
  eig_matrices(g, E, G);
  cx ra = g.rp, rb = g.rsv;
  cx k[4];  // kmpu, kmsu (from yb), km1pd, km1sd (from yt)
  for (int r = 0; r < 4; r++) {
    const double* y = (r < 2) ? yb : yt;
    k[r] = G[r][0] * y[0] + G[r][1] * y[1] + G[r][2] * y[2] + G[r][3] * y[3];
  }
  const int I[6] = {0, 0, 1, 1, 2, 3}, J[6] = {0, 2, 1, 3, 2, 3};
  if (half) {
    for (int t = 0; t < 6; t++) {
      int i = I[t], j = J[t];
      cx v = E[i][2] * E[j][2] * k[2] * k[2] / (2.0 * ra) +
             (E[i][2] * E[j][3] + E[i][3] * E[j][2]) * k[2] * k[3] / (ra + rb) +
             E[i][3] * E[j][3] * k[3] * k[3] / (2.0 * rb);
      out[t] = v.r;
    }
    return;
  }
  cx FA = ffunc(ra, dm), GA = gfunc(ra, dm), FB = ffunc(rb, dm), GB = gfunc(rb, dm);
  cx H1 = h1func(ra, rb, dm), H2 = h2func(ra, rb, dm);
  for (int t = 0; t < 6; t++) {
    int i = I[t], j = J[t];
    cx v = E[i][0] * E[j][0] * k[0] * k[0] * FA + E[i][2] * E[j][2] * k[2] * k[2] * FA +
           E[i][1] * E[j][1] * k[1] * k[1] * FB + E[i][3] * E[j][3] * k[3] * k[3] * FB +
           H1 * ((E[i][0] * E[j][1] + E[i][1] * E[j][0]) * k[0] * k[1] +
                 (E[i][2] * E[j][3] + E[i][3] * E[j][2]) * k[2] * k[3]) +
           H2 * ((E[i][0] * E[j][3] + E[i][3] * E[j][0]) * k[0] * k[3] +
                 (E[i][1] * E[j][2] + E[i][2] * E[j][1]) * k[2] * k[1]) +
           GA * (E[i][0] * E[j][2] + E[i][2] * E[j][0]) * k[0] * k[2] +
           GB * (E[i][1] * E[j][3] + E[i][3] * E[j][1]) * k[1] * k[3];
    out[t] = v.r;
  }
}

DK_HD double cnorm5(cx e[5]) {
  double t = 0.0;
  for (int i = 0; i < 5; i++) t = fmax(t, cabs_(e[i]));
  if (t < 1.0e-40) t = 1.0;
  for (int i = 0; i < 5; i++) e[i] = e[i] / t;
  return log(t);
}
DK_HD double rnorm4(double e[4]) {
  double t = 0.0;
  for (int i = 0; i < 4; i++) t = fmax(t, fabs(e[i]));
  if (t < 1.0e-40) t = 1.0;
  for (int i = 0; i < 4; i++) e[i] = e[i] / t;
  return log(t);
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
// Scratch: per layer 10 doubles of up-sweep compound vector + 1 exponent, 4 of
// interface eigenfunction and 4 of unnormalised kernel factors; element
// (m, q) lives at scratch[(m * 19 + q) * stride].
// emit(m, dcdah, dcdbv, dcdn) receives, top layer first, tregn96's kernels after
// the sphericity back-mapping, rounded to float as chksiz does.  lay(m) returns
// the flattened Layer, vtpf(m) the float sphericity factor.
template <class LayerFn, class VtpFn, class EmitFn>
DK_HD int solve(int L, LayerFn lay, VtpFn vtpf, double omega, double c, double* scratch,
                long stride, EmitFn emit) {
  auto S = [&](int m, int q) -> double& { return scratch[(long)(m * 19 + q) * stride]; };
  const double wvno = omega / c, wvno2 = wvno * wvno, om2 = omega * omega;
  Eig g;
  Hyp h;
  // ---- up: compound constraint vectors
  cx cd[5];
  eigen(lay(L - 1), omega, wvno, &g);
  halfspace_constraint(g, cd);
  for (int q = 0; q < 5; q++) {
    S(L - 1, 2 * q) = cd[q].r;
    S(L - 1, 2 * q + 1) = cd[q].i;
  }
  S(L - 1, 10) = 0.0;
  double exsum = 0.0;
  for (int m = L - 2; m >= 0; m--) {
    const Layer ly = lay(m);
    eigen(ly, omega, wvno, &g);
    hyper(g, ly.d, &h);
    cx CA[5][5];
    compound(g, h, CA);
    cx ee[5];
    for (int i = 0; i < 5; i++) {
      cx cr = mk(0.0);
      for (int j = 0; j < 5; j++) cr = cr + cd[j] * CA[j][i];
      ee[i] = cr;
    }
    double exn = cnorm5(ee);
    exsum = exsum + h.pex + h.svex + exn;
    for (int q = 0; q < 5; q++) {
      cd[q] = ee[q];
      S(m, 2 * q) = ee[q].r;
      S(m, 2 * q + 1) = ee[q].i;
    }
    S(m, 10) = exsum;
  }
  // ---- surface eigenfunction
  const double exe1 = S(0, 10);
  cx c0_2 = mk(S(0, 2), S(0, 3)), c0_3 = mk(S(0, 4), S(0, 5));
  const double f1213 = -S(0, 2);
  if (!(f1213 != 0.0) || !isfinite(f1213)) return -1;
  S(0, 11) = (c0_3 / c0_2).r;
  S(0, 12) = 1.0;
  S(0, 13) = 0.0;
  S(0, 14) = 0.0;
  // ---- down sweep, interface eigenfunctions and layer integrals
  double vv[4] = {1.0, 0.0, 0.0, 0.0};
  double exd = 0.0;
  double sumi0 = 0, sumi1 = 0, sumi2 = 0, sumi3 = 0;
  for (int m = 0; m < L; m++) {
    const Layer ly = lay(m);
    eigen(ly, omega, wvno, &g);
    if (m < L - 1) {
      hyper(g, ly.d, &h);
      double sP = 1.0, sS = 1.0, cpex;
      if (h.pex > h.svex) {
        sS = (h.pex - h.svex > 40.0) ? 0.0 : exp(-(h.pex - h.svex));
        cpex = h.pex;
      } else {
        sP = (h.svex - h.pex > 40.0) ? 0.0 : exp(-(h.svex - h.pex));
        cpex = h.svex;
      }
      double AA[4][4];
      propagator(g, h, sP, sS, AA);
      double a0[4];
      for (int i = 0; i < 4; i++)
        a0[i] = AA[i][0] * vv[0] + AA[i][1] * vv[1] + AA[i][2] * vv[2] + AA[i][3] * vv[3];
      double ex2 = rnorm4(a0);
      exd = exd + cpex + ex2;
      for (int i = 0; i < 4; i++) vv[i] = a0[i];
      // contraction at interface m+1 (svfunc)
      const int n = m + 1;
      double cd1 = S(n, 0), cd2 = S(n, 2), cd3 = S(n, 4), cd4 = -S(n, 4), cd5 = S(n, 6),
             cd6 = S(n, 8);
      double tz1 = -vv[3], tz2 = -vv[2], tz3 = vv[1], tz4 = vv[0];
      double uu1 = tz2 * cd6 - tz3 * cd5 + tz4 * cd4;
      double uu2 = -tz1 * cd6 + tz3 * cd3 - tz4 * cd2;
      double uu3 = tz1 * cd5 - tz2 * cd3 + tz4 * cd1;
      double uu4 = -tz1 * cd4 + tz2 * cd2 - tz3 * cd1;
      double ext = exd + S(n, 10) - exe1;
      double y[4] = {0, 0, 0, 0};
      if (ext > -80.0 && ext < 80.0) {
        double fact = exp(ext);
        y[0] = uu1 * fact / f1213;
        y[1] = uu2 * fact / f1213;
        y[2] = uu3 * fact / f1213;
        y[3] = uu4 * fact / f1213;
      }
      for (int i = 0; i < 4; i++) S(n, 11 + i) = y[i];
    }
    double yt[4], yb[4];
    for (int i = 0; i < 4; i++) {
      yt[i] = S(m, 11 + i);
      yb[i] = (m < L - 1) ? S(m + 1, 11 + i) : 0.0;
    }
    double I[6];
    layer_integrals(g, ly.d, yt, yb, m == L - 1, I);
    const double INT11 = I[0], INT13 = I[1], INT22 = I[2], INT24 = I[3], INT33 = I[4],
                 INT44 = I[5];
    // energy (getmat coefficients)
    const double rho = ly.rho;
    const double ah = sqrt(ly.A / rho), av = sqrt(ly.C / rho), bv = sqrt(ly.L / rho);
    const double eta = ly.F / (ly.A - 2.0 * ly.L);
    const double a12 = -wvno, a14 = 1.0 / ly.L, a21 = wvno * ly.F / ly.C, a23 = 1.0 / ly.C;
    const double URUR = INT11, UZUZ = INT22;
    const double DURDUR = a12 * a12 * INT22 + 2.0 * a12 * a14 * INT24 + a14 * a14 * INT44;
    const double DUZDUZ = a21 * a21 * INT11 + 2.0 * a21 * a23 * INT13 + a23 * a23 * INT33;
    const double URDUZ = a21 * INT11 + a23 * INT13;
    const double UZDUR = a12 * INT22 + a14 * INT24;
    sumi0 = sumi0 + rho * (URUR + UZUZ);
    sumi1 = sumi1 + ly.L * UZUZ + ly.A * URUR;
    sumi2 = sumi2 + ly.L * UZDUR - ly.F * URDUZ;
    sumi3 = sumi3 + ly.L * DURDUR + ly.C * DUZDUZ;
    S(m, 15) = rho * ah * (URUR - 2.0 * eta * URDUZ / wvno);                          // facah
    S(m, 16) = rho * av * DUZDUZ / wvno2;                                             // facav
    S(m, 17) = rho * bv * (UZUZ + 2.0 * UZDUR / wvno + DURDUR / wvno2 + 4.0 * eta * URDUZ / wvno);
    S(m, 18) = -ly.F * URDUZ / (wvno * eta);                                          // facn
  }
  const double ugr = (wvno * sumi1 + sumi2) / (omega * sumi0);
  const double nrm = ugr * sumi0;
  // ---- Q dispersion of c (gammap: Qp = 150, Qs = 50, fref = 1 Hz), used only by
  //      the sphericity factor tm below
  const double pi = 3.141592653589493;
  const double qai = (double)(1.0f / 150.0f), qbi = (double)(1.0f / 50.0f);
  const double lw = log(omega / (2.0 * pi * 1.0)) / pi;
  double dc = 0.0;
  for (int m = 0; m < L; m++) {
    const Layer ly = lay(m);
    double ah = sqrt(ly.A / ly.rho), av = sqrt(ly.C / ly.rho), bv = sqrt(ly.L / ly.rho);
    double xs = (S(m, 17) / nrm) * bv * qbi;
    double xp = (S(m, 16) / nrm) * av * qai + (S(m, 15) / nrm) * ah * qai;
    dc = dc + lw * xs + lw * xp;
  }
  const double cq = omega / wvno + dc;
  const double ar = 6370.0;
  const double tm = sqrt(1.0 + (cq / (2.0 * ar * omega)) * (cq / (2.0 * ar * omega)));
  const double tm3 = tm * tm * tm;
  (void)om2;
  auto f32 = [](double x) -> float { return fabs(x) < 1.0e-36 ? 0.0f : (float)x; };
  for (int m = 0; m < L; m++) {
    const double vt = (double)vtpf(m);
    emit(m, f32(S(m, 15) / nrm * vt / tm3), f32(S(m, 17) / nrm * vt / tm3), f32(S(m, 18) / nrm));
  }
  return 0;
}

// depthkernelTI splice in float32 (depthkernelTI.f90:152-162): interval j sums
// its nsub[j] sublayers of dc/dA*TA + dc/dL*TL.
DK_HD float splice_term(float rrho, float rvp, float rvs, float TA, float TL, float TF,
                        float dah, float dbv, float dn) {
#pragma clang fp contract(off)
  float den = TA - 2.0f * TL;
  float den2 = den * den;
  float dA = 0.5f / (rrho * rvp) * dah - TF / den2 * dn;
  float dL = 0.5f / (rrho * rvs) * dbv + 2.0f * TF / den2 * dn;
  return dA * TA + dL * TL;
}

}  // namespace ti
}  // namespace dk
