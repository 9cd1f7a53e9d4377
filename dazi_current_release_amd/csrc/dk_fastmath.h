// FP64 elementary functions of the stage's layer algebra, written for gfx950:
// the arguments they take are bounded by the physics (a layer's phase
// p = |nu| d, an exponent sum), which lets them drop the general library's
// range handling.  Host builds (CPU self-checks) use the same formulas with
// the host's fma and sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#ifndef DK_HD
#define DK_HD __host__ __device__ inline
#endif

namespace dk {

// a*b + c as one three-operand v_fma_f64 (the backend's two-operand
// v_fmac_f64 form would first copy a loop-invariant addend into the result
// register, one extra move per Horner step).  fma3 takes the constant addend
// c from a VGPR pair, fma3s from an SGPR pair (the one scalar operand a VALU
// instruction may read): SGPR constants free vector registers, at the cost of
// scalar moves where the scalar registers run out.  The full-precision exp and
// sin/cos (kernel, VJP and TI passes, which sit at their two-wave register
// limit) take every coefficient from SGPRs; the root search's shorter sin/cos
// keeps its coefficients in VGPRs.
DK_HD double fma3(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}
DK_HD double fma3s(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

// exp(-p) for p >= -700 (0 once it underflows), for the secular function,
// its derivatives and the TI eigenfunction scales: Cody-Waite reduction p = n ln2 - r with the two-part ln2 of
// fdlibm, r in [-ln2/2, ln2/2], and the degree-11 Taylor polynomial of e^r
// (remainder below 7e-15 relative), scaled by 2^-n.
DK_HD double exp_neg(double p) {
  p = fmin(p, 745.0);  // beyond, e^-p is below the smallest denormal
  const double n = rint(p * 1.4426950408889634074);
  double r = fma(n, 6.93147180369123816490e-01, -p);
  r = fma(n, 1.90821492927058770002e-10, r);
  double q = 2.5052108385441718775e-08;              // 1/11!
  q = fma3s(q, r, 2.7557319223985890653e-07);          // 1/10!
  q = fma3s(q, r, 2.7557319223985892511e-06);          // 1/9!
  q = fma3s(q, r, 2.4801587301587301566e-05);          // 1/8!
  q = fma3s(q, r, 1.9841269841269841253e-04);          // 1/7!
  q = fma3s(q, r, 1.3888888888888889419e-03);          // 1/6!
  q = fma3s(q, r, 8.3333333333333332177e-03);          // 1/5!
  q = fma3s(q, r, 4.1666666666666664354e-02);          // 1/4!
  q = fma3s(q, r, 1.6666666666666665741e-01);          // 1/3!
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, -(int)n);
}

// sin and cos of p, 0 <= p below about 1e5 (a layer's phase p = sqrt(-X) d is
// at most omega d / beta): p = n pi/2 + r with the
// three-part pi/2 of fdlibm (Cody-Waite), r in [-pi/4, pi/4], Taylor
// polynomials to degree 15 and 16 (remainders below 3e-14 and 1e-15), and the
// quadrant n mod 4.
DK_HD void sincos_phase(double p, double* sp, double* cp) {
  const double n = rint(p * 6.36619772367581382433e-01);
  double r = fma(n, -1.57079632673412561417e+00, p);
  r = fma(n, -6.07710050650619224932e-11, r);
  r = fma(n, -2.02226624879595063154e-21, r);
  const double r2 = r * r;
  double ps = 7.6471637318198164759e-13;               // 1/15!
  ps = fma3s(ps, r2, -1.6059043836821614599e-10);        // -1/13!
  ps = fma3s(ps, r2, 2.5052108385441718775e-08);         // 1/11!
  ps = fma3s(ps, r2, -2.7557319223985892511e-06);        // -1/9!
  ps = fma3s(ps, r2, 1.9841269841269841253e-04);         // 1/7!
  ps = fma3s(ps, r2, -8.3333333333333332177e-03);        // -1/5!
  ps = fma3s(ps, r2, 1.6666666666666665741e-01);         // 1/3!, subtracted below
  const double sn = fma(-r * r2, ps, r);
  double pc = 4.7794773323873852974e-14;               // 1/16!
  pc = fma3s(pc, r2, -1.1470745597729724714e-11);        // -1/14!
  pc = fma3s(pc, r2, 2.0876756987868098979e-09);         // 1/12!
  pc = fma3s(pc, r2, -2.7557319223985890653e-07);        // -1/10!
  pc = fma3s(pc, r2, 2.4801587301587301566e-05);         // 1/8!
  pc = fma3s(pc, r2, -1.3888888888888889419e-03);        // -1/6!
  pc = fma3s(pc, r2, 4.1666666666666664354e-02);         // 1/4!
  pc = fma(pc, r2, -0.5);
  const double cs = fma(pc, r2, 1.0);
  const int q = (int)n & 3;
  const double s0 = (q & 1) ? cs : sn, c0 = (q & 1) ? sn : cs;
  *sp = (q & 2) ? -s0 : s0;
  *cp = ((q + 1) & 2) ? -c0 : c0;
}

// Shorter forms for the root search, which needs F's sign and its root to far
// better than the 2e-6 km/s refinement bracket but not to full precision.
// Their polynomials are near-minimax fits on the reduced intervals, computed
// for this code by Lawson's iteratively reweighted least squares at 3,000
// Chebyshev nodes (low-order terms fixed to their Taylor values); the errors
// quoted are the maxima measured against long-double libm over 400,001 points.
//  * exp_neg_rs: e^r = 1 + r + r^2/2 + r^3 q(r), q of degree 5 (1.3e-12
//    relative on |r| <= ln2/2), same reduction as exp_neg; no clamp of p: a
//    layer's phase is bounded by omega d / c (a few thousand at most), where
//    ldexp already gives 0.
//  * sincos_phase_rs: p = n pi + r, |r| <= pi/2 (the three-part pi of fdlibm,
//    doubled), sin to degree 11 (1.7e-11 absolute) and cos to degree 12
//    (1.2e-12 absolute), and (-1)^n as a flip of both sign bits: fewer
//    instructions than the quadrant selects of a reduction by pi/2.
// An error e in every layer's wave functions moves a root by about e times
// the c-scale of F's variation (a few 1e-11 km/s).
DK_HD double exp_neg_rs(double p) {
  const double n = rint(p * 1.4426950408889634074);
  double r = fma(n, 6.93147180369123816490e-01, -p);
  r = fma(n, 1.90821492927058770002e-10, r);
  double q = 2.4750213656569838e-05;
  q = fma3s(q, r, 1.9907280660463785e-04);
  q = fma3s(q, r, 1.3889063765214424e-03);
  q = fma3s(q, r, 8.333284946702695e-03);
  q = fma3s(q, r, 4.1666665753399536e-02);
  q = fma3s(q, r, 1.6666666770246064e-01);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, -(int)n);
}

DK_HD void sincos_phase_rs(double p, double* sp, double* cp) {
  const double n = rint(p * 3.18309886183790671538e-01);  // 1/pi
  double r = fma(n, -3.14159265346825122834e+00, p);
  r = fma(n, -1.21542010130123844986e-10, r);
  r = fma(n, -4.04453249759190126308e-21, r);
  const double r2 = r * r;
  double ps = 2.3846694416234846e-08;
  ps = fma3(ps, r2, -2.752261887781079e-06);
  ps = fma3(ps, r2, 1.9840804039690193e-04);
  ps = fma3(ps, r2, -8.333330495675608e-03);
  ps = fma3(ps, r2, 1.6666666606467112e-01);  // subtracted below
  const double sn = fma(-r * r2, ps, r);
  double pc = 1.997495792156871e-09;
  pc = fma3(pc, r2, -2.752987105362138e-07);
  pc = fma3(pc, r2, 2.4801191757009406e-05);
  pc = fma3(pc, r2, -1.3888886240518484e-03);
  pc = fma3(pc, r2, 4.1666666602470245e-02);
  pc = fma(pc, r2, -0.5);
  const double cs = fma(pc, r2, 1.0);
  // (-1)^n: n's parity into both sign bits
  const unsigned long long flip = (unsigned long long)(((unsigned)(int)n & 1u) << 31) << 32;
  *sp = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, sn) ^ flip);
  *cp = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, cs) ^ flip);
}

// 1/sqrt(x) for x > 0: the hardware estimate and one Newton step.
DK_HD double rsqrt_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
  return y * fma(-0.5 * x * y, y, 1.5);
#else
  return 1.0 / sqrt(x);
#endif
}

}  // namespace dk
