// Host build of the root search's secular function (dk_fdm.h), for the CPU
// tests that compare it with the oracle's dltar4 (surfdisp96.f:767-865) on the
// same flattened models: test infrastructure, not part of libdazi_dk.so and
// not called by anything the library runs.
#include <vector>

#include "dk_fdm.h"

// F(c) at `period` for L flattened layers (float32, half-space last) as
// mantissa, with the power-of-two exponent in *ex; the records are built as
// k_forward builds them in LDS.
extern "C" double dk_host_secular(double c, double period, int L, const float* d, const float* a,
                                  const float* b, const float* rho, int* ex) {
  std::vector<double4> mdl(2 * (size_t)L);
  for (int m = 0; m < L; m++) {
    const double4 l = make_double4(d[m], a[m], b[m], rho[m]);
    const int u = m > 0 ? m - 1 : 0;
    const double4 up = make_double4(d[u], a[u], b[u], rho[u]);
    dk::layer_rec(l, up, m == 0, &mdl[2 * m], &mdl[2 * m + 1]);
  }
  const double omega = 2.0 * 3.141592653589793 / period;
  return dk::secular_fwd(c, omega, L, mdl.data(), ex);
}

// Kernels -(dF/dm)/(dF/dc) at (c, period) in the flattened (a, b, rho, d) of
// each layer (out: 4 L), from the building blocks the kernel pass and the VJP
// use; returns dF/dc (mantissa, exponent in *ex).
extern "C" double dk_host_kernels(double c, double period, int L, const float* d, const float* a,
                                  const float* b, const float* rho, double* out, int* ex) {
  if (L < 1 || L > 1024) return 0.0;
  std::vector<double4> lay((size_t)L);
  for (int m = 0; m < L; m++) lay[m] = make_double4(d[m], a[m], b[m], rho[m]);
  return dk::fdm_kernels_ref(c, 2.0 * 3.141592653589793 / period, L, lay.data(), out, ex);
}

// The root search's shorter elementary functions (dk_fastmath.h), for the CPU
// test of their error bounds.
extern "C" double dk_host_exp_neg_rs(double p) { return dk::exp_neg_rs(p); }
extern "C" void dk_host_sincos_phase_rs(double p, double* s, double* c) {
  dk::sincos_phase_rs(p, s, c);
}
