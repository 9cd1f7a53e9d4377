// Period azimuthal-anisotropy map: the forward prediction of the 2-psi
// Rayleigh anisotropy from Lsen_Gsc and the Gc, Gs layer models
// (code/src_inv_iso_joint/FwdAzimuthalAniMap.f90:44-89).  One thread per
// (interior cell, period):
//   cosRc = sum_k Lsen_Gsc(post, t, k) * gcf(ii, jj, k)    (float32, k order)
//   sinRc = sum_k Lsen_Gsc(post, t, k) * gsf(ii, jj, k)
//   AzimAmp     = 0.5 * sqrt(cosRc^2 + sinRc^2)
//   AzimAmpRela = AzimAmp / real(tRcV(cell, t))
//   AzimAng     = 0.5 * (atan2(sinRc, cosRc) / pi * 180, + 360 when negative)
// with the reference's float32 operations (no contraction; sqrt and the
// division evaluated in float64 and rounded once, which is the correctly
// rounded float32 result), its pi (a default-real literal, 3.14159274), and
// atan2 in float64 rounded to float32 (the reference's libm atan2f differs
// from it by one unit in the last place for some inputs).  Reads are
// coalesced along the cells of a row: the map is bound by HBM bytes.
#include <hip/hip_runtime.h>

#include "dk_internal.h"

namespace dk {

__global__ void k_azimuth_map(int nvx, int nvz, int nd, int K, const float* __restrict__ lsen,
                              const float* __restrict__ gcf, const float* __restrict__ gsf,
                              const double* __restrict__ tRcV, float* __restrict__ cos_rc,
                              float* __restrict__ sin_rc, float* __restrict__ amp,
                              float* __restrict__ amp_rela, float* __restrict__ angle) {
#pragma clang fp contract(off)
  const int cells = nvx * nvz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)cells * K) return;
  const int cell = (int)(t % cells), tt = (int)(t / cells);
  const int jj = cell / nvx, ii = cell - jj * nvx;
  const long ncol = (long)(nvx + 2) * (nvz + 2);
  const long post = (long)(jj + 1) * (nvx + 2) + ii + 1;
  float c = 0.0f, s = 0.0f;
  for (int k = 0; k < nd; k++) {
    const float l = lsen[((long)k * K + tt) * ncol + post];
    c = c + l * gcf[(long)k * cells + cell];
    s = s + l * gsf[(long)k * cells + cell];
  }
  cos_rc[t] = c;
  sin_rc[t] = s;
  const float a = 0.5f * (float)sqrt((double)(c * c + s * s));
  if (amp) amp[t] = a;
  if (amp_rela) amp_rela[t] = (float)((double)a / (double)(float)tRcV[t]);
  if (angle) {
    const double pi = (double)3.1415926535898f;
    float g = (float)((double)(float)atan2((double)s, (double)c) / pi * 180.0);
    if (g < 0.0f) g = g + 360.0f;
    angle[t] = 0.5f * g;
  }
}

hipError_t launch_azimuth_map(int nvx, int nvz, int nz, int K, const float* lsen, const float* gcf,
                              const float* gsf, const double* tRcV, float* cos_rc, float* sin_rc,
                              float* amp, float* amp_rela, float* angle, hipStream_t st) {
  const long n = (long)nvx * nvz * K;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_azimuth_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nvx, nvz,
                     nz - 1, K, lsen, gcf, gsf, tRcV, cos_rc, sin_rc, amp, amp_rela, angle);
  return hipGetLastError();
}

}  // namespace dk
