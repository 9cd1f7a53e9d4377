// Per-layer geometry of a node model that every column shares: the parts of
// the isotropic earth flattening (sphere, surfdisp96.f:480-547) and of the TI
// flattening and sphericity factors (sphere_tdisp96 and bldsph,
// tregn96.f:774-841, :1301-1383) that depend on the layer thicknesses only,
// and the sublayer mapping of the node refinement (refineLayerMdl,
// CalSurfGAniso_Joint.f90:146-196).  k_prep_geom writes one record per refined
// layer; k_prep_nodes and k_prep_ti apply them to every column.
#pragma once
#include <hip/hip_runtime.h>

#include "dk_internal.h"

namespace dk {

// TI part: rho's factor tmp^-2.275 and the moduli's tmp^-0.275 of the TI
// flattening (radius 6371), the layer thickness, its flattened value and the
// float32 sphericity factor (radius 6370).
struct TiGeom {
  double pr, fm;
  float d, df, vtp, pad;
};

// Isotropic part: radii r0 (top) and r1 = 6370 - dr (bottom, dr the thickness
// sum down to the layer's bottom), the flattened thickness, the velocity factor
// tmp and its float32 value bt, rho's factor bt^-2.275 in float32 (the
// flattening) and in float64 (the chain rule of the kernel passes).
struct FlatGeom {
  double tmp, btr;
  float d, bt, rpow;
  // node models (k_prep_geom): the layer lies in node interval `node` as
  // sublayer `sub` (1-based) of `ns`; the half-space takes node `node` itself
  // (sub = ns = 0); node = -1 marks a layer the sublayer rule never forms.
  // bad (half-space record only): the rule formed other than L - 1 layers.
  int node, sub, ns, bad;
  TiGeom ti;
};
static_assert(sizeof(FlatGeom) == PREP_GEOM_BYTES, "prep geometry record");

__device__ __forceinline__ FlatGeom flat_geom(double r0, double dr, bool half) {
#pragma clang fp contract(off)
  const double ar = 6370.0;
  const double r1 = ar - dr;
  const double z0 = ar * log(ar / r0);
  const double z1 = ar * log(ar / r1);
  FlatGeom g = {};
  g.d = half ? 0.0f : (float)(z1 - z0);
  g.tmp = (ar + ar) / (r0 + r1);
  g.bt = (float)g.tmp;
  g.rpow = powf(g.bt, -2.275f);
  g.btr = pow((double)g.bt, -2.275);
  return g;
}

// State of the TI recurrences, pushed top to bottom one layer at a time.
struct TiFlat {
  double r0 = 6371.0;  // flattening recurrence
  double rv = 6370.0;  // sphericity recurrence
};

__device__ __forceinline__ TiGeom ti_geom(TiFlat& s, float d, bool half) {
#pragma clang fp contract(off)
  const double ar = 6371.0;
  const double e_r = (double)(-2.275f), e_m = (double)(-0.2750f);
  TiGeom g = {};
  float td = half ? 1.0f : d;
  double r1 = s.r0 - (double)td;
  double z0 = ar * log(ar / s.r0);
  double z1 = ar * log(ar / r1);
  float tdf = (float)(z1 - z0);
  double tmp = (ar + ar) / (s.r0 + r1);
  g.pr = pow(tmp, e_r);
  g.fm = pow(tmp, e_m);
  g.d = d;
  g.df = half ? 0.0f : tdf;
  s.r0 = r1;
  const double ar2 = 6370.0;
  double zd = half ? 1.0 : (double)g.df;
  double rv1 = s.rv * exp(-zd / ar2);
  double tv = (ar2 + ar2) / (s.rv + rv1);
  g.vtp = (float)tv;
  s.rv = rv1;
  return g;
}

}  // namespace dk
