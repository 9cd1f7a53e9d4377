/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's
 * per-column dispersion + FD depth-kernel path. It is the checker for the HIP
 * path and the "port" CPU baseline; no product code links or calls it.
 *
 * Follows, statement by statement where precision matters:
 *   code/src_inv_iso_joint/surfdisp96.f:52-354   surfdisp96 driver (Rayleigh, mode 1, phase)
 *   code/src_inv_iso_joint/surfdisp96.f:361-382  gtsolh   (real*4 Newton on the half-space equation)
 *   code/src_inv_iso_joint/surfdisp96.f:384-476  getsol   (bracket scan, direction, floor, failure)
 *   code/src_inv_iso_joint/surfdisp96.f:480-547  sphere   (earth flattening, ar = 6370)
 *   code/src_inv_iso_joint/surfdisp96.f:551-680  nevill + half (Neville / bisection refinement)
 *   code/src_inv_iso_joint/surfdisp96.f:767-865  dltar4   (Dunkin compound-matrix secular function)
 *   code/src_inv_iso_joint/surfdisp96.f:868-985  var
 *   code/src_inv_iso_joint/surfdisp96.f:989-1014 normc
 *   code/src_inv_iso_joint/surfdisp96.f:1018-1062 dnka
 *   code/src_inv_iso_joint/DepthKernel1D.f90:102-200 depthkernel_column_fd + Brocher
 *   code/src_inv_iso_joint/CalSurfG.f90:2490-2549 refineGrid2LayerMdl
 *   code/src_inv_iso_joint/CalSurfG.f90:86-107    OpenMP column loop of depthkernel
 *
 * Precision follows the reference: the layered model, the flattening output,
 * gtsolh and the FD perturbations are float32; the secular function and the
 * root search are float64; c is rounded to float32 on output (sngl). Compile
 * with -ffp-contract=off and without -ffast-math so that float32 expressions
 * round exactly as the reference's x86-64 build does.
 *
 * Two knobs exist only for building tight oracles (tests of analytic kernels):
 * tol_rel replaces the 1e-6 relative convergence test of nevill, and
 * round_f32 = 0 keeps c in float64 on output.  The reference values are
 * tol_rel = 1e-6, round_f32 = 1.
 *
 * The reference keeps `del1st` in a SAVE variable shared by all OpenMP
 * threads; here it lives in a per-call state, which is what a serial run of
 * the reference computes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NL 200
#define ORC_NP 160

typedef struct {
  double del1st;
  double tol_rel;
  int round_f32;
  int nevals;      /* secular-function evaluations (diagnostics) */
} orc_state;

/* ---------------------------------------------------------------- gtsolh */
float orc_gtsolh(float a, float b) {
  float c = 0.95f * b;
  for (int i = 0; i < 5; i++) {
    float gamma = b / a;
    float kappa = c / b;
    float k2 = kappa * kappa;
    float gk = gamma * kappa;
    float gk2 = gk * gk;
    float fac1 = sqrtf(1.0f - gk2);
    float fac2 = sqrtf(1.0f - k2);
    float fr = (2.0f - k2) * (2.0f - k2) - 4.0f * fac1 * fac2;
    float frp = -(4.0f * (2.0f - k2) * kappa) + 4.0f * fac2 * gamma * gamma * kappa / fac1 +
                4.0f * fac1 * kappa / fac2;
    frp = frp / b;
    c = c - fr / frp;
  }
  return c;
}

/* ---------------------------------------------------------------- var */
static void orc_var(double p, double q, double ra, double rb, double wvno, double xka,
                    double xkb, double dpth, double* w_o, double* cosp_o, double* exa_o,
                    double* a0_o, double* cpcq, double* cpy, double* cpz, double* cqw,
                    double* cqx, double* xy, double* xz, double* wy, double* wz) {
  double exa = 0.0, a0 = 1.0, pex = 0.0, sex = 0.0;
  double sinp, w = 0, x = 0, cosp = 0, sinq, y = 0, z = 0, cosq = 0, fac;
  if (wvno < xka) {
    sinp = sin(p);
    w = sinp / ra;
    x = -ra * sinp;
    cosp = cos(p);
  } else if (wvno == xka) {
    cosp = 1.0;
    w = dpth;
    x = 0.0;
  } else if (wvno > xka) {
    pex = p;
    fac = 0.0;
    if (p < 16) fac = exp(-2.0 * p);
    cosp = (1.0 + fac) * 0.5;
    sinp = (1.0 - fac) * 0.5;
    w = sinp / ra;
    x = ra * sinp;
  }
  if (wvno < xkb) {
    sinq = sin(q);
    y = sinq / rb;
    z = -rb * sinq;
    cosq = cos(q);
  } else if (wvno == xkb) {
    cosq = 1.0;
    y = dpth;
    z = 0.0;
  } else if (wvno > xkb) {
    sex = q;
    fac = 0.0;
    if (q < 16) fac = exp(-2.0 * q);
    cosq = (1.0 + fac) * 0.5;
    sinq = (1.0 - fac) * 0.5;
    y = sinq / rb;
    z = rb * sinq;
  }
  exa = pex + sex;
  a0 = 0.0;
  if (exa < 60.0) a0 = exp(-exa);
  *cpcq = cosp * cosq;
  *cpy = cosp * y;
  *cpz = cosp * z;
  *cqw = cosq * w;
  *cqx = cosq * x;
  *xy = x * y;
  *xz = x * z;
  *wy = w * y;
  *wz = w * z;
  /* surfdisp96.f:978-983 rescales cosq, y, z after this point; the products
     above are already formed, so the rescaled values are never used here. */
  *w_o = w;
  *cosp_o = cosp;
  *exa_o = exa;
  *a0_o = a0;
}

/* ---------------------------------------------------------------- dnka */
static void orc_dnka(double ca[5][5], double wvno2, double gam, double gammk, double rho,
                     double a0, double cpcq, double cpy, double cpz, double cqw, double cqx,
                     double xy, double xz, double wy, double wz) {
  const double one = 1.0, two = 2.0;
  double gamm1 = gam - one;
  double twgm1 = gam + gamm1;
  double gmgmk = gam * gammk;
  double gmgm1 = gam * gamm1;
  double gm1sq = gamm1 * gamm1;
  double rho2 = rho * rho;
  double a0pq = a0 - cpcq;
  /* ca[i][j] is the reference's ca(i+1, j+1) */
  ca[0][0] = cpcq - two * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
  ca[0][1] = (wvno2 * cpy - cqx) / rho;
  ca[0][2] = -(twgm1 * a0pq + gammk * xz + wvno2 * gamm1 * wy) / rho;
  ca[0][3] = (cpz - wvno2 * cqw) / rho;
  ca[0][4] = -(two * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) / rho2;
  ca[1][0] = (gmgmk * cpz - gm1sq * cqw) * rho;
  ca[1][1] = cpcq;
  ca[1][2] = gammk * cpz - gamm1 * cqw;
  ca[1][3] = -wz;
  ca[1][4] = ca[0][3];
  ca[3][0] = (gm1sq * cpy - gmgmk * cqx) * rho;
  ca[3][1] = -xy;
  ca[3][2] = gamm1 * cpy - gammk * cqx;
  ca[3][3] = ca[1][1];
  ca[3][4] = ca[0][1];
  ca[4][0] = -(two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2;
  ca[4][1] = ca[3][0];
  ca[4][2] = -(gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * xz + gamm1 * gm1sq * wy) * rho;
  ca[4][3] = ca[1][0];
  ca[4][4] = ca[0][0];
  double t = -two * wvno2;
  ca[2][0] = t * ca[4][2];
  ca[2][1] = t * ca[3][2];
  ca[2][2] = a0 + two * (cpcq - ca[0][0]);
  ca[2][3] = t * ca[1][2];
  ca[2][4] = t * ca[0][2];
}

/* ---------------------------------------------------------------- normc */
static void orc_normc(double ee[5]) {
  double t1 = 0.0;
  for (int i = 0; i < 5; i++)
    if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
  if (t1 < 1.e-40) t1 = 1.0;
  for (int i = 0; i < 5; i++) ee[i] = ee[i] / t1;
}

/* ---------------------------------------------------------------- dltar4 */
double orc_dltar4(double wvno, double omga, const float* d, const float* a, const float* b,
                  const float* rho, int mmax, int llw) {
  double e[5], ee[5], ca[5][5];
  double omega = omga;
  if (omega < 1.0e-4) omega = 1.0e-4;
  double wvno2 = wvno * wvno;
  double xka = omega / (double)a[mmax - 1];
  double xkb = omega / (double)b[mmax - 1];
  double wvnop = wvno + xka;
  double wvnom = fabs(wvno - xka);
  double ra = sqrt(wvnop * wvnom);
  wvnop = wvno + xkb;
  wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double t = (double)b[mmax - 1] / omega;
  double gammk = 2.0 * t * t;
  double gam = gammk * wvno2;
  double gamm1 = gam - 1.0;
  double rho1 = (double)rho[mmax - 1];
  e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -rho1 * ra;
  e[2] = rho1 * (gamm1 - gammk * ra * rb);
  e[3] = rho1 * rb;
  e[4] = wvno2 - ra * rb;
  for (int m = mmax - 2; m >= llw - 1; m--) {
    xka = omega / (double)a[m];
    xkb = omega / (double)b[m];
    t = (double)b[m] / omega;
    gammk = 2.0 * t * t;
    gam = gammk * wvno2;
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    double dpth = (double)d[m];
    rho1 = (double)rho[m];
    double p = ra * dpth;
    double q = rb * dpth;
    double w, cosp, exa, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
    orc_var(p, q, ra, rb, wvno, xka, xkb, dpth, &w, &cosp, &exa, &a0, &cpcq, &cpy, &cpz, &cqw,
            &cqx, &xy, &xz, &wy, &wz);
    orc_dnka(ca, wvno2, gam, gammk, rho1, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz);
    for (int i = 0; i < 5; i++) {
      double cr = 0.0;
      for (int j = 0; j < 5; j++) cr = cr + e[j] * ca[j][i];
      ee[i] = cr;
    }
    orc_normc(ee);
    for (int i = 0; i < 5; i++) e[i] = ee[i];
  }
  if (llw != 1) {
    /* water layer on top (surfdisp96.f:844-860) */
    xka = omega / (double)a[0];
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    double dpth = (double)d[0];
    rho1 = (double)rho[0];
    double p = ra * dpth;
    double znul = 1.0e-05;
    double w, cosp, exa, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
    orc_var(p, znul, ra, znul, wvno, xka, znul, dpth, &w, &cosp, &exa, &a0, &cpcq, &cpy, &cpz,
            &cqw, &cqx, &xy, &xz, &wy, &wz);
    double w0 = -rho1 * w;
    return cosp * e[0] + w0 * e[1];
  }
  return e[0];
}

/* ---------------------------------------------------------------- sphere */
typedef struct {
  float rtp[ORC_NL], dtp[ORC_NL], btp[ORC_NL];
  float dhalf;
} orc_sph;

static void orc_sphere(int ifunc, int iflag, float* d, float* a, float* b, float* rho,
                       orc_sph* s, int mmax) {
  double ar = 6370.0, dr = 0.0, r0 = ar, r1, z0, z1, tmp;
  d[mmax - 1] = 1.0f;
  if (iflag == 0) {
    for (int i = 0; i < mmax; i++) {
      s->dtp[i] = d[i];
      s->rtp[i] = rho[i];
    }
    for (int i = 0; i < mmax; i++) {
      dr = dr + (double)d[i];
      r1 = ar - dr;
      z0 = ar * log(ar / r0);
      z1 = ar * log(ar / r1);
      d[i] = (float)(z1 - z0);
      tmp = (ar + ar) / (r0 + r1);
      a[i] = (float)((double)a[i] * tmp);
      b[i] = (float)((double)b[i] * tmp);
      s->btp[i] = (float)tmp;
      r0 = r1;
    }
    s->dhalf = d[mmax - 1];
  } else {
    d[mmax - 1] = s->dhalf;
    for (int i = 0; i < mmax; i++) {
      if (ifunc == 1)
        rho[i] = s->rtp[i] * powf(s->btp[i], -5.0f);
      else if (ifunc == 2)
        rho[i] = s->rtp[i] * powf(s->btp[i], -2.275f);
    }
  }
  d[mmax - 1] = 0.0f;
}

/* ---------------------------------------------------------------- half / nevill */
typedef struct {
  const float *d, *a, *b, *rho;
  int mmax, llw;
  orc_state* st;
} orc_mdl;

static double orc_F(const orc_mdl* M, double wvno, double omega) {
  M->st->nevals++;
  return orc_dltar4(wvno, omega, M->d, M->a, M->b, M->rho, M->mmax, M->llw);
}

static double dsign1(double x) { return copysign(1.0, x); }

static void orc_half(double c1, double c2, double* c3, double* del3, double omega,
                     const orc_mdl* M) {
  *c3 = 0.5 * (c1 + c2);
  double wvno = omega / *c3;
  *del3 = orc_F(M, wvno, omega);
}

static double orc_nevill(double t, double c1, double c2, double del1, double del2,
                         const orc_mdl* M, double twopi) {
  double x[20], y[20];
  double c3, del3, s13, s32, ss1, ss2, s1, s2, denom, wvno;
  int nev, nctrl, m = 1;
  const double f001 = (double)0.01f; /* real*4 literal 0.01 promoted */
  double omega = twopi / t;
  orc_half(c1, c2, &c3, &del3, omega, M);
  nev = 1;
  nctrl = 1;
  for (;;) {
    nctrl = nctrl + 1;
    if (nctrl >= 100) break;
    if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
      nev = 0;
      orc_half(c1, c2, &c3, &del3, omega, M);
    }
    s13 = del1 - del3;
    s32 = del3 - del2;
    if (dsign1(del3) * dsign1(del1) < 0.0) {
      c2 = c3;
      del2 = del3;
    } else {
      c1 = c3;
      del1 = del3;
    }
    if (fabs(c1 - c2) <= M->st->tol_rel * c1) break;
    if (dsign1(s13) != dsign1(s32)) nev = 0;
    ss1 = fabs(del1);
    s1 = f001 * ss1;
    ss2 = fabs(del2);
    s2 = f001 * ss2;
    if (s1 > ss2 || s2 > ss1 || nev == 0) {
      orc_half(c1, c2, &c3, &del3, omega, M);
      nev = 1;
      m = 1;
    } else {
      if (nev == 2) {
        x[m] = c3;
        y[m] = del3;
      } else {
        x[0] = c1;
        y[0] = del1;
        x[1] = c2;
        y[1] = del2;
        m = 1;
      }
      int bail = 0;
      for (int kk = 1; kk <= m; kk++) {
        int j = m - kk + 1; /* 1-based */
        denom = y[m] - y[j - 1];
        if (fabs(denom) < 1.0e-10 * fabs(y[m])) {
          bail = 1;
          break;
        }
        x[j - 1] = (-y[j - 1] * x[j] + y[m] * x[j - 1]) / denom;
      }
      if (!bail) {
        c3 = x[0];
        wvno = omega / c3;
        del3 = orc_F(M, wvno, omega);
        nev = 2;
        m = m + 1;
        if (m > 10) m = 10;
      } else {
        orc_half(c1, c2, &c3, &del3, omega, M);
        nev = 1;
        m = 1;
      }
    }
  }
  return c3;
}

/* ---------------------------------------------------------------- getsol */
static int orc_getsol(double t1, double* c1p, double clow, double dc, double cm, float betmx,
                      int ifirst, const orc_mdl* M) {
  double twopi = 2.0 * 3.141592653589793;
  double c1 = *c1p, c2, del1, del2, plmn, omega, wvno;
  int idir;
  omega = twopi / t1;
  wvno = omega / c1;
  del1 = orc_F(M, wvno, omega);
  if (ifirst == 1) M->st->del1st = del1;
  plmn = dsign1(M->st->del1st) * dsign1(del1);
  if (ifirst == 1)
    idir = +1;
  else if (plmn >= 0.0)
    idir = +1;
  else
    idir = -1;
  for (;;) {
    if (idir > 0)
      c2 = c1 + dc;
    else
      c2 = c1 - dc;
    if (c2 <= clow) {
      idir = +1;
      c1 = clow;
      continue;
    }
    omega = twopi / t1;
    wvno = omega / c2;
    del2 = orc_F(M, wvno, omega);
    if (dsign1(del1) != dsign1(del2)) {
      double cn = orc_nevill(t1, c1, c2, del1, del2, M, twopi);
      c1 = cn;
      *c1p = c1;
      if (c1 > (double)betmx) return -1;
      return 1;
    }
    c1 = c2;
    del1 = del2;
    *c1p = c1;
    if (c1 < cm) return -1;
    if (c1 >= ((double)betmx + dc)) return -1;
  }
}

/* ---------------------------------------------------------------- surfdisp96
 * Rayleigh (iwave=2), fundamental mode, phase velocity (igr=0).
 * Returns the 1-based index of the first failed period (0 when all found);
 * cg(k..kmax) are zero-filled from that period on, as the reference does.
 * If flat_out is non-NULL it receives the flattened model (4*nlayer floats:
 * d, a, b, rho) exactly as dltar4 sees it.
 */
int orc_surfdisp96_x(const float* thkm, const float* vpm, const float* vsm, const float* rhom,
                     int nlayer, int iflsph, int kmax, const double* t, double* cg,
                     double tol_rel, int round_f32, float* flat_out, int* nevals_out) {
  float d[ORC_NL], a[ORC_NL], b[ORC_NL], rho[ORC_NL];
  orc_sph sph;
  orc_state st;
  st.del1st = 0.0;
  st.tol_rel = tol_rel;
  st.round_f32 = round_f32;
  st.nevals = 0;
  double c[ORC_NP];
  int mmax = nlayer;
  if (mmax > ORC_NL || kmax > ORC_NP || mmax < 1) return -2;
  for (int i = 0; i < mmax; i++) {
    b[i] = vsm[i];
    a[i] = vpm[i];
    d[i] = thkm[i];
    rho[i] = rhom[i];
  }
  int llw = 1;
  if (b[0] <= 0.0f) llw = 2;
  if (iflsph == 1) orc_sphere(0, 0, d, a, b, rho, &sph, mmax);
  int jmn = 0, jsol = 1;
  float betmx = -1.e20f, betmn = 1.e20f;
  for (int i = 0; i < mmax; i++) {
    if (b[i] > 0.01f && b[i] < betmn) {
      betmn = b[i];
      jmn = i;
      jsol = 1;
    } else if (b[i] <= 0.01f && a[i] < betmn) {
      betmn = a[i];
      jmn = i;
      jsol = 0;
    }
    if (b[i] > betmx) betmx = b[i];
  }
  if (iflsph == 1) orc_sphere(2, 1, d, a, b, rho, &sph, mmax);
  if (flat_out) {
    for (int i = 0; i < mmax; i++) {
      flat_out[i] = d[i];
      flat_out[mmax + i] = a[i];
      flat_out[2 * mmax + i] = b[i];
      flat_out[3 * mmax + i] = rho[i];
    }
  }
  const float ddc = 0.005f;
  const double onea = (double)1.5f;
  float cc1;
  if (jsol == 0)
    cc1 = betmn;
  else
    cc1 = orc_gtsolh(a[jmn], b[jmn]);
  cc1 = 0.95f * cc1;
  cc1 = 0.90f * cc1;
  double cc = (double)cc1;
  double dc = fabs((double)ddc);
  double c1 = cc, clow = cc, cm = cc;
  for (int i = 0; i < kmax; i++) c[i] = 0.0;
  orc_mdl M = {d, a, b, rho, mmax, llw, &st};
  int failed = 0;
  for (int k = 0; k < kmax; k++) {
    double t1 = t[k];
    int ifirst;
    if (k == 0) {
      c1 = cc;
      clow = cc;
      ifirst = 1;
    } else {
      ifirst = 0;
      c1 = c[k - 1] - onea * dc;
      clow = cm;
    }
    int iret = orc_getsol(t1, &c1, clow, dc, cm, betmx, ifirst, &M);
    if (iret == -1) {
      failed = k + 1;
      break;
    }
    c[k] = c1;
    cg[k] = round_f32 ? (double)(float)c[k] : c[k];
  }
  if (failed)
    for (int i = failed - 1; i < kmax; i++) cg[i] = 0.0;
  if (nevals_out) *nevals_out = st.nevals;
  return failed;
}

int orc_surfdisp96(const float* thkm, const float* vpm, const float* vsm, const float* rhom,
                   int nlayer, int iflsph, int kmax, const double* t, double* cg) {
  return orc_surfdisp96_x(thkm, vpm, vsm, rhom, nlayer, iflsph, kmax, t, cg, 1.0e-6, 1, NULL,
                          NULL);
}

/* ---------------------------------------------------------------- Brocher + refinement */
/* DepthKernel1D.f90:187-200. The reference compiler expands x**n for integer
   n as ((x*x)*x)*x..., so the powers are formed by sequential products. */
void orc_brocher(int nz, const float* vs, float* vp, float* rho) {
  for (int k = 0; k < nz; k++) {
    float v = vs[k];
    float v2 = v * v, v3 = v2 * v, v4 = v3 * v;
    vp[k] = 0.9409f + 2.0947f * v - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
    float p = vp[k];
    float p2 = p * p, p3 = p2 * p, p4 = p3 * p, p5 = p4 * p;
    rho[k] = 1.6612f * p - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
  }
}

/* refineGrid2LayerMdl / refineLayerMdl; returns rmax (number of refined layers
   including the half-space), or -1 when NL would be exceeded. */
int orc_refine(float minthk0, int mmax, const float* dep, const float* vp, const float* vs,
               const float* rho, float* rvp, float* rvs, float* rrho, float* rthk,
               int* nsublay) {
  int k = 0;
  for (int i = 0; i < mmax - 1; i++) {
    float thk = dep[i + 1] - dep[i];
    float minthk = thk / minthk0;
    int ns = (int)((thk + 1.0e-4f) / minthk) + 1;
    if (nsublay) nsublay[i] = ns;
    float newthk = thk / (float)ns;
    for (int j = 1; j <= ns; j++) {
      if (k >= ORC_NL - 1) return -1;
      rthk[k] = newthk;
      float w = (float)(2 * j - 1);
      float den = (float)(2 * ns);
      rvp[k] = vp[i] + w * (vp[i + 1] - vp[i]) / den;
      rvs[k] = vs[i] + w * (vs[i + 1] - vs[i]) / den;
      rrho[k] = rho[i] + w * (rho[i + 1] - rho[i]) / den;
      k++;
    }
  }
  rthk[k] = 0.0f;
  rvp[k] = vp[mmax - 1];
  rvs[k] = vs[mmax - 1];
  rrho[k] = rho[mmax - 1];
  return k + 1;
}

/* ---------------------------------------------------------------- FD column kernel
 * depthkernel_column_fd with an optional explicit vp/rho (vp_in/rho_in non-NULL);
 * with both NULL, vp and rho come from Brocher(vs) as in the reference.
 * sen_* are (kmax, nz) column-major: sen[i*kmax + k].
 */
int orc_column_fd(int nz, const float* vs, const float* vp_in, const float* rho_in, int kmax,
                  const double* periods, const float* depz, float minthk, float dlnvs,
                  float dlnvp, float dlnrho, double tol_rel, int round_f32, int exact_step,
                  double* pv, double* sen_vs, double* sen_vp, double* sen_rho) {
  float vp[ORC_NL], rho[ORC_NL];
  float vsm[ORC_NL], vpm[ORC_NL], rhom[ORC_NL];
  float rvp[ORC_NL], rvs[ORC_NL], rrho[ORC_NL], rthk[ORC_NL];
  double cg1[ORC_NP], cg2[ORC_NP];
  if (kmax > ORC_NP || nz > ORC_NL) return -2;
  if (vp_in && rho_in) {
    for (int i = 0; i < nz; i++) {
      vp[i] = vp_in[i];
      rho[i] = rho_in[i];
    }
  } else {
    orc_brocher(nz, vs, vp, rho);
  }
  int rmax = orc_refine(minthk, nz, depz, vp, vs, rho, rvp, rvs, rrho, rthk, NULL);
  if (rmax < 0) return -2;
  int fail0 = orc_surfdisp96_x(rthk, rvp, rvs, rrho, rmax, 1, kmax, periods, pv, tol_rel,
                               round_f32, NULL, NULL);
  if (!sen_vs) return fail0;
  for (int i = 0; i < nz; i++) {
    vsm[i] = vs[i];
    vpm[i] = vp[i];
    rhom[i] = rho[i];
  }
  for (int i = 0; i < nz; i++) {
    float* arrs[3] = {vsm, vpm, rhom};
    const float* base[3] = {vs, vp, rho};
    float dln[3] = {dlnvs, dlnvp, dlnrho};
    double* outs[3] = {sen_vs, sen_vp, sen_rho};
    for (int p = 0; p < 3; p++) {
      float h = 0.5f * dln[p] * base[p][i];
      float vlo = base[p][i] - h, vhi = base[p][i] + h;
      arrs[p][i] = vlo;
      rmax = orc_refine(minthk, nz, depz, vpm, vsm, rhom, rvp, rvs, rrho, rthk, NULL);
      orc_surfdisp96_x(rthk, rvp, rvs, rrho, rmax, 1, kmax, periods, cg1, tol_rel, round_f32,
                       NULL, NULL);
      arrs[p][i] = base[p][i] + h;
      rmax = orc_refine(minthk, nz, depz, vpm, vsm, rhom, rvp, rvs, rrho, rthk, NULL);
      orc_surfdisp96_x(rthk, rvp, rvs, rrho, rmax, 1, kmax, periods, cg2, tol_rel, round_f32,
                       NULL, NULL);
      arrs[p][i] = base[p][i];
      /* reference: (cg2-cg1)/(dln*v); exact_step divides by the float32 step
         actually taken (tight oracles only) */
      double den = exact_step ? ((double)vhi - (double)vlo) : (double)(dln[p] * base[p][i]);
      for (int nn = 0; nn < kmax; nn++) outs[p][i * kmax + nn] = (cg2[nn] - cg1[nn]) / den;
    }
  }
  return fail0;
}

/* ---------------------------------------------------------------- depthkernel stage
 * The reference's depthkernel (CalSurfG.f90:1-113, default FD backend): OpenMP
 * over columns, outputs in the caller's Fortran layout
 *   pv(nx*ny, kmax), sen_*(nx*ny, kmax, nz), column index fastest.
 * vel is (nx, ny, nz) Fortran order. nthreads <= 0 keeps the OpenMP default.
 * Pass sen_vs == NULL to compute phase velocities only.
 */
void orc_depthkernel(int nx, int ny, int nz, const float* vel, const float* vp_opt,
                     const float* rho_opt, double* pv, double* sen_vs, double* sen_vp,
                     double* sen_rho, int kmax, const double* tRc, const float* depz,
                     float minthk, int nthreads) {
  const float dln = 0.01f;
  long ncol = (long)nx * ny;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (long post = 0; post < ncol; post++) {
    float vsz[ORC_NL], vpz[ORC_NL], rhoz[ORC_NL];
    double pvc[ORC_NP];
    double* s1 = NULL, *s2 = NULL, *s3 = NULL;
    if (sen_vs) {
      s1 = (double*)malloc(sizeof(double) * 3 * kmax * nz);
      s2 = s1 + kmax * nz;
      s3 = s2 + kmax * nz;
    }
    for (int k = 0; k < nz; k++) {
      vsz[k] = vel[post + (long)k * ncol];
      if (vp_opt) vpz[k] = vp_opt[post + (long)k * ncol];
      if (rho_opt) rhoz[k] = rho_opt[post + (long)k * ncol];
    }
    orc_column_fd(nz, vsz, vp_opt ? vpz : NULL, rho_opt ? rhoz : NULL, kmax, tRc, depz, minthk,
                  dln, dln, dln, 1.0e-6, 1, 0, pvc, s1, s2, s3);
    for (int k = 0; k < kmax; k++) pv[post + (long)k * ncol] = pvc[k];
    if (sen_vs) {
      for (int i = 0; i < nz; i++)
        for (int k = 0; k < kmax; k++) {
          long o = post + ncol * ((long)k + (long)kmax * i);
          sen_vs[o] = s1[i * kmax + k];
          sen_vp[o] = s2[i * kmax + k];
          sen_rho[o] = s3[i * kmax + k];
        }
      free(s1);
    }
  }
}

/* Layered-model forward only, for ragged batches: column j has nl[j] layers at
 * offset loff[j] in the flat arrays and kk[j] periods at offset toff[j].
 * tol_rel and round_f32 as orc_surfdisp96_x (1e-6 and 1: the reference's). */
void orc_forward_layered_x(int ncol, const int* nl, const long* loff, const int* kk,
                           const long* toff, const float* thk, const float* vp, const float* vs,
                           const float* rho, const double* periods, double* cg, int* fail,
                           int nthreads, double tol_rel, int round_f32) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int j = 0; j < ncol; j++) {
    int f = orc_surfdisp96_x(thk + loff[j], vp + loff[j], vs + loff[j], rho + loff[j], nl[j], 1,
                             kk[j], periods + toff[j], cg + toff[j], tol_rel, round_f32, NULL,
                             NULL);
    if (fail) fail[j] = f;
  }
}

void orc_forward_layered(int ncol, const int* nl, const long* loff, const int* kk,
                         const long* toff, const float* thk, const float* vp, const float* vs,
                         const float* rho, const double* periods, double* cg, int* fail,
                         int nthreads) {
  orc_forward_layered_x(ncol, nl, loff, kk, toff, thk, vp, vs, rho, periods, cg, fail, nthreads,
                        1.0e-6, 1);
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
