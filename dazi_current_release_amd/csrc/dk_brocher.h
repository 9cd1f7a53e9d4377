// Brocher (2005) coefficients of the G builder's Vs row, in float32 as the
// reference evaluates them (CalSurfGAniso_Joint.f90:754-760, real arithmetic,
// left to right): coe_a = dVp/dVs at the node Vs, coe_rho = dRho/dVp * coe_a.
#pragma once

#include <hip/hip_runtime.h>

namespace dk {

__host__ __device__ inline void brocher_coeffs(float v, float* coe_a, float* coe_rho) {
#pragma clang fp contract(off)
  const float v2 = v * v, v3 = v2 * v, v4 = v3 * v;
  const float a = 2.0947f - 0.8206f * 2.0f * v + 0.2683f * 3.0f * v2 - 0.0251f * 4.0f * v3;
  const float p = 0.9409f + 2.0947f * v - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
  const float p2 = p * p, p3 = p2 * p, p4 = p3 * p;
  *coe_a = a;
  *coe_rho = a * (1.6612f - 0.4721f * 2.0f * p + 0.0671f * 3.0f * p2 - 0.0043f * 4.0f * p3 +
                  0.000106f * 5.0f * p4);
}

// eff = (sen_vp coe_a + sen_rho coe_rho) + sen_vs in float64 with separate
// roundings, the reference's mixed real*8/real*4 expression
// (CalSurfGAniso_Joint.f90:762-765).
__host__ __device__ inline double effective_vs(double svs, double svp, double srho, float coe_a,
                                               float coe_rho) {
#pragma clang fp contract(off)
  return (svp * (double)coe_a + srho * (double)coe_rho) + svs;
}

}  // namespace dk
