// HIP kernels of the depth-kernel stage for gfx950 (CDNA4).
//
//   k_prep_nodes   node model (vel, depz, minthk) -> refined, earth-flattened
//                  layered model + start velocity; one wavefront per column.
//                  (CalSurfG.f90:2490-2549, DepthKernel1D.f90:187-200,
//                   surfdisp96.f:99-207 with sphere :480-547, gtsolh :361-382)
//   k_prep_layers  the same for layered (ragged) models.
//   k_order_*      the scheduling order of a ragged batch (a counting sort of
//                  its columns and solves by layer count, then period count).
//   k_forward<S>   Rayleigh fundamental-mode root search: S lanes of a wave own
//                  one column and walk its periods in order, reproducing the
//                  reference's scan grid, direction, floor and failure rules
//                  (surfdisp96.f:223-348, getsol :384-476); each scan step
//                  evaluates S trial velocities at once (the secular
//                  function of dk_fdm.h) and a __ballot finds the first sign
//                  change; the root is verified in a window around an
//                  inverse-cubic estimate, else refined by S-point
//                  multisection, and placed by a final secant step.
//   k_kernels      per (column, period) implicit differentiation of the
//                  secular function at the root, dc/dm = -(dF/dm)/(dF/dc):
//                  a bottom-up sweep with the tangent in c and checkpoints,
//                  and a hand-written adjoint sweep of the potential-basis
//                  recursion (dk_fdm.h) over segments recomputed into LDS;
//                  maps refined-layer partials to the node or layer outputs.
//   k_vjp          the same sweeps per (column, period), with a per-layer
//                  tape, contracted with grad_c over a column's periods into
//                  layer gradients (two columns per wavefront).
//
// All secular-function arithmetic is FP64 on the vector ALUs; nothing here is
// a dense contraction, so MFMA is not used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dk_brocher.h"
#include "dk_fdm.h"
#include "dk_geom.h"
#include "dk_internal.h"

namespace dk {

// ------------------------------------------------------------------ float32-exact model prep
// Evaluated with contraction off so that every float32 and float64 operation
// rounds as the reference's separate multiply/add instructions do.
__device__ float brocher_vp(float v) {
#pragma clang fp contract(off)
  float v2 = v * v, v3 = v2 * v, v4 = v3 * v;
  return 0.9409f + 2.0947f * v - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
}
__device__ float brocher_rho(float p) {
#pragma clang fp contract(off)
  float p2 = p * p, p3 = p2 * p, p4 = p3 * p, p5 = p4 * p;
  return 1.6612f * p - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
}

__device__ float gtsolh_f32(float a, float b) {
#pragma clang fp contract(off)
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

// One layer (thk, vp, vs, rho as float32 values) flattened with its geometry
// into (d, a, b, rho) rounded to float32 as the reference stores them; *ok = 0
// for a non-finite or non-positive value (the half-space may have d = 0).
__device__ __forceinline__ double4 flat_layer(double4 l, bool half, const FlatGeom& g, bool* ok) {
#pragma clang fp contract(off)
  const float d = g.d;
  float a = (float)((double)(float)l.y * g.tmp);
  float b = (float)((double)(float)l.z * g.tmp);
  float rho = (float)l.w * g.rpow;
  const double4 o = make_double4(d, a, b, rho);
  *ok = isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(o.w) && o.y > 0.0 &&
        o.z > 0.0 && o.w > 0.0 && (o.x > 0.0 || half);
  return o;
}

// Flattening geometry and node mapping of the refined layers of a node model,
// which every column shares (refineLayerMdl, CalSurfGAniso_Joint.f90:146-196,
// then sphere and the TI factors): thread m runs the sublayer rule down to
// layer m and sums the thicknesses in the order a per-column pass would, so
// every value is the one a per-column pass (and ti_geom's recurrence) computes.
__global__ void k_prep_geom(int nz, const float* __restrict__ depz, float minthk0, int L,
                            FlatGeom* __restrict__ geo, float* __restrict__ geo_depz) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L) return;
  if (m < nz) geo_depz[m] = depz[m];  // L >= nz: every depth is copied
  double dr = 0.0, dr0 = 0.0;
  int k = 0, node = -1, sub = 0, nsm = 0;
  for (int i = 0; i < nz - 1 && k <= m; i++) {
    float thk = depz[i + 1] - depz[i];
    float minthk = thk / minthk0;
    int ns = (int)((thk + 1.0e-4f) / minthk) + 1;
    float newthk = thk / (float)ns;
    for (int j = 1; j <= ns && k < L - 1 && k <= m; j++, k++) {
      dr0 = dr;
      dr = dr + (double)newthk;
      if (k == m) {
        node = i;
        sub = j;
        nsm = ns;
      }
    }
  }
  const bool half = (m == L - 1);
  if (half) {  // the half-space is flattened with a unit thickness
    dr0 = dr;
    dr = dr + 1.0;
  }
  FlatGeom g = flat_geom(6370.0 - dr0, dr, half);
  g.node = half ? nz - 1 : node;
  g.sub = sub;
  g.ns = nsm;
  g.bad = half ? (k != L - 1) : 0;
  // the TI recurrences run through every layer above m (the sphericity chain
  // needs each one's flattened thickness): the same sequence a per-column pass
  // of the TI prep makes, once per geometry
  TiFlat ts;
  k = 0;
  for (int i = 0; i < nz - 1 && k <= m; i++) {
    float thk = depz[i + 1] - depz[i];
    float minthk = thk / minthk0;
    int ns = (int)((thk + 1.0e-4f) / minthk) + 1;
    float newthk = thk / (float)ns;
    for (int j = 1; j <= ns && k < L - 1 && k <= m; j++, k++) {
      const TiGeom t = ti_geom(ts, newthk, false);
      if (k == m) g.ti = t;
    }
  }
  if (half) g.ti = ti_geom(ts, 0.0f, true);
  geo[m] = g;
}

// Node models: vel is (ncol, nz) column-fastest (Fortran vel(nx,ny,nz)).
// vp/rho optional (both or neither); otherwise Brocher from vs.  geo holds
// k_prep_geom's records of the same depz, minthk0 and L.  One wavefront per
// column, lane = layer: each lane interpolates its layer between its two
// nodes and flattens it; the start-velocity layer (the first layer of least
// b, or a where b <= 0.01: the layer surfdisp96's sequential rule keeps),
// betmx and the bad flags are wave reductions.
__global__ void __launch_bounds__(256) k_prep_nodes(
    int ncol, int nz, const float* __restrict__ vel, const float* __restrict__ vpin,
    const float* __restrict__ rhoin, const float* __restrict__ depz,
    const float* __restrict__ geo_depz, int L, const FlatGeom* __restrict__ geo,
    double4* __restrict__ layers, double* __restrict__ btp, double* __restrict__ btr,
    double* __restrict__ cc, float* __restrict__ betmx, int* __restrict__ bad_out) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  // the records were built from geo_depz: a depz with other values (a graph
  // replayed after depz changed) must not use them
  bool moved = false;
  for (int i = lane; i < nz; i += 64)
    moved = moved || (__float_as_uint(depz[i]) != __float_as_uint(geo_depz[i]));
  moved = __ballot(moved) != 0ull;
  auto node_vals = [&](int col, int i, float* vs, float* vp, float* rho) {
    const size_t o = (size_t)i * ncol + col;
    *vs = vel[o];
    if (vpin) {
      *vp = vpin[o];
      *rho = rhoin[o];
    } else {
      *vp = brocher_vp(*vs);
      *rho = brocher_rho(*vp);
    }
  };
  for (int col = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < ncol; col += nw) {
    double4* lay = layers + (size_t)col * L;
    float kmin = __builtin_inff(), amn = 0.0f, bmn = 0.0f, bmx = -1.e20f;
    bool have = false, bad = false;
    for (int m0 = 0; m0 < L; m0 += 64) {
      const int m = m0 + lane;
      bool cand = false, isbad = false;
      float key = __builtin_inff(), a = 0.0f, b = -1.e20f;
      if (m < L) {
        const FlatGeom g = geo[m];
        if (g.node >= 0) {
          const bool half = (m == L - 1);
          float vs, vp, rho;
          node_vals(col, g.node, &vs, &vp, &rho);
          if (!half) {
            float vs_n, vp_n, rho_n;
            node_vals(col, g.node + 1, &vs_n, &vp_n, &rho_n);
            const float den = (float)(2 * g.ns);
            const float w = (float)(2 * g.sub - 1);
            vp = vp + w * (vp_n - vp) / den;
            vs = vs + w * (vs_n - vs) / den;
            rho = rho + w * (rho_n - rho) / den;
          }
          bool ok;
          const double4 o = flat_layer(make_double4(0.0, vp, vs, rho), half, g, &ok);
          lay[m] = o;
          btp[(size_t)col * L + m] = g.bt;
          btr[(size_t)col * L + m] = g.btr;
          a = (float)o.y;
          b = (float)o.z;
          isbad = !ok || (m == 0 && o.z <= 0.0) || g.bad;
          if (b > 0.01f) {
            key = b;
            cand = true;
          } else if (b <= 0.01f) {
            key = a;
            cand = true;
          }
          cand = cand && (key < 1.e20f);
          if (!cand) key = __builtin_inff();
        }
      }
      if (m0 == 0) {  // the rule starts from layer 0's values
        amn = __shfl(a, 0, 64);
        bmn = __shfl(b, 0, 64);
      }
      float kw = key, bw = b;
#pragma unroll
      for (int sh = 1; sh < 64; sh <<= 1) {
        kw = fminf(kw, __shfl_xor(kw, sh, 64));
        bw = fmaxf(bw, __shfl_xor(bw, sh, 64));
      }
      bmx = fmaxf(bmx, bw);
      bad = bad || (__ballot(isbad) != 0ull);
      const uint64_t at = __ballot(cand && key == kw);
      if (at != 0ull && (!have || kw < kmin)) {
        const int j = __ffsll((unsigned long long)at) - 1;
        kmin = kw;
        amn = __shfl(a, j, 64);
        bmn = __shfl(b, j, 64);
        have = true;
      }
    }
    // flat_finish: jsol = 0 (a fluid start layer) is outside this stage's model class
    const bool jsol0 = have && !(bmn > 0.01f);
    float cc1 = gtsolh_f32(amn, bmn);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    if (!(cc1 > 0.0f) || !isfinite(cc1)) bad = true;
    if (lane == 0) {
      cc[col] = (double)cc1;
      betmx[col] = bmx;
      bad_out[col] = moved ? DK_PREP_GEOMETRY : (bad || jsol0) ? DK_PREP_BADMODEL : 0;
    }
  }
}

// Float64 flattening of one layer (the differentiable path): radii r0 = 6370 -
// (thickness sum above the layer) and r1 = 6370 - (the sum down to its bottom).
__device__ __forceinline__ double4 flat_layer_f64(double4 l, double r0, double r1, bool half,
                                                  double* tmp_out, double* tr_out) {
  const double ar = 6370.0;
  const double d = half ? 0.0 : ar * log(ar / r1) - ar * log(ar / r0);
  const double tmp = (ar + ar) / (r0 + r1);
  const double tr = pow(tmp, -2.275);
  *tmp_out = tmp;
  *tr_out = tr;
  return make_double4(d, l.y * tmp, l.z * tmp, l.w * tr);
}

__device__ __forceinline__ double readlane_f64(double v, int j) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Layered models (ragged): thk/vp/vs/rho concatenated at loff[col], float32
// (reference semantics) or, when thk64 is set, float64 arrays flattened in
// float64 (sphere, surfdisp96.f:480-547, then the start-velocity rule,
// :139-207).  One wavefront per column, lane = layer, 64 layers at a time:
// lane m adds the thicknesses of layers 0..m one by one in layer order (the
// sequential recurrence's sums, bit for bit), flattens its layer, and the
// start-velocity layer (the first layer of least b, or a where b <= 0.01),
// betmx and the bad flags are wave reductions, as in k_prep_nodes.
__global__ void __launch_bounds__(256) k_prep_layers(
    int ncol, const int* __restrict__ Lc, const int64_t* __restrict__ loff,
    const float* __restrict__ thk, const float* __restrict__ vp, const float* __restrict__ vs,
    const float* __restrict__ rho, const double* __restrict__ thk64,
    const double* __restrict__ vp64, const double* __restrict__ vs64,
    const double* __restrict__ rho64, double4* __restrict__ layers, double* __restrict__ btp,
    double* __restrict__ btr, double* __restrict__ cc, float* __restrict__ betmx,
    int* __restrict__ bad_out) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  const bool f64 = thk64 != nullptr;
  for (int col = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < ncol; col += nw) {
    const int L = Lc[col];
    const int64_t o = loff[col];
    if (L < 1) {
      if (lane == 0) bad_out[col] = 1;
      continue;
    }
    double carry = 0.0;  // thickness sum of the layers of the previous chunks
    float kmin = __builtin_inff(), amn = 0.0f, bmn = 0.0f, bmx = -1.e20f;
    bool have = false, bad = false;
    for (int m0 = 0; m0 < L; m0 += 64) {
      const int m = m0 + lane;
      const bool in = m < L, half = (m == L - 1);
      double4 l = make_double4(0.0, 0.0, 0.0, 0.0);
      if (in)
        l = f64 ? make_double4(thk64[o + m], vp64[o + m], vs64[o + m], rho64[o + m])
                : make_double4(thk[o + m], vp[o + m], vs[o + m], rho[o + m]);
      // the half-space enters the sum with a unit thickness
      const double t = !in ? 0.0 : f64 ? (half ? 1.0 : l.x) : (double)(half ? 1.0f : (float)l.x);
      const int n = min(64, L - m0);
      double acc = carry, accp = carry;
      for (int jj = 0; jj < n; jj++) {
        const double tj = readlane_f64(t, jj);
        if (jj == lane) accp = acc;
        if (jj <= lane) acc = acc + tj;
      }
      carry = readlane_f64(acc, n - 1);
      bool cand = false, isbad = false;
      float key = __builtin_inff(), a = 0.0f, b = -1.e20f;
      if (in) {
        double4 f;
        double bt, br;
        bool ok;
        if (f64) {
          f = flat_layer_f64(l, 6370.0 - accp, 6370.0 - acc, half, &bt, &br);
          ok = isfinite(f.x) && isfinite(f.y) && isfinite(f.z) && isfinite(f.w) && f.y > 0.0 &&
               f.z > 0.0 && f.w > 0.0 && (f.x > 0.0 || half);
        } else {
          const FlatGeom g = flat_geom(6370.0 - accp, acc, half);
          f = flat_layer(l, half, g, &ok);
          bt = g.bt;
          br = g.btr;
        }
        layers[o + m] = f;
        btp[o + m] = bt;
        btr[o + m] = br;
        a = (float)f.y;
        b = (float)f.z;
        isbad = !ok || (m == 0 && f.z <= 0.0);
        if (b > 0.01f) {
          key = b;
          cand = true;
        } else if (b <= 0.01f) {
          key = a;
          cand = true;
        }
        cand = cand && (key < 1.e20f);
        if (!cand) key = __builtin_inff();
      }
      if (m0 == 0) {  // the rule starts from layer 0's values
        amn = __shfl(a, 0, 64);
        bmn = __shfl(b, 0, 64);
      }
      float kw = key, bw = b;
#pragma unroll
      for (int sh = 1; sh < 64; sh <<= 1) {
        kw = fminf(kw, __shfl_xor(kw, sh, 64));
        bw = fmaxf(bw, __shfl_xor(bw, sh, 64));
      }
      bmx = fmaxf(bmx, bw);
      bad = bad || (__ballot(isbad) != 0ull);
      const uint64_t at = __ballot(cand && key == kw);
      if (at != 0ull && (!have || kw < kmin)) {
        const int jm = __ffsll((unsigned long long)at) - 1;
        kmin = kw;
        amn = __shfl(a, jm, 64);
        bmn = __shfl(b, jm, 64);
        have = true;
      }
    }
    // a fluid start layer (jsol = 0) is outside this stage's model class
    const bool jsol0 = have && !(bmn > 0.01f);
    float cc1 = gtsolh_f32(amn, bmn);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    if (!(cc1 > 0.0f) || !isfinite(cc1)) bad = true;
    if (lane == 0) {
      cc[col] = (double)cc1;
      betmx[col] = bmx;
      bad_out[col] = (bad || jsol0) ? DK_PREP_BADMODEL : 0;
    }
  }
}

// ------------------------------------------------------------------ forward root search
enum { ST_START = 0, ST_UP = 1, ST_DN = 2, ST_REF = 3, ST_VER = 4, ST_DONE = 5 };
#define DK_MAX_REFINE 48
#define DK_MAX_ROUNDS 100000

__device__ __forceinline__ void col_geom(const BatchDev& B, int col, int* L, int64_t* loff, int* K,
                                         int64_t* toff) {
  if (B.uniform) {
    *L = B.Lu;
    *loff = (int64_t)col * B.Lu;
    *K = B.Ku;
    *toff = 0;
  } else {
    *L = B.L[col];
    *loff = B.loff[col];
    *K = B.K[col];
    *toff = B.toff[col];
  }
}
__device__ __forceinline__ int64_t solve_index(const BatchDev& B, int col, int64_t toff, int k) {
  return B.uniform ? (int64_t)k * B.ncol + col : toff + k;
}
__device__ __forceinline__ double period_of(const BatchDev& B, int64_t toff, int k) {
  return B.uniform ? B.periods[k] : B.periods[toff + k];
}

template <int S>
__device__ __forceinline__ uint64_t group_bits(uint64_t mask, int gbase) {
  if (S == 64) return mask;
  return (mask >> gbase) & ((1ull << S) - 1ull);
}
__device__ __forceinline__ int ffs64(uint64_t m) { return m ? __ffsll((unsigned long long)m) - 1 : -1; }

// Value-and-exponent helpers for the bracket ends.
struct FV {
  double m;
  int e;
};
__device__ __forceinline__ double fv_ratio_root(double lo, double hi, FV flo, FV fhi) {
  // root of the chord through (lo, flo) and (hi, fhi), flo and fhi of opposite sign
  int eb = max(flo.e, fhi.e);
  double a = ldexp(flo.m, flo.e - eb);
  double b = ldexp(fhi.m, fhi.e - eb);
  double den = a - b;
  if (!(fabs(den) > 0.0)) return 0.5 * (lo + hi);
  double t = a / den;
  if (!(t >= 0.0 && t <= 1.0)) return 0.5 * (lo + hi);
  return lo + (hi - lo) * t;
}

// Root at y = 0 of the cubic x(y) through the points (X[t], Y[t]) with ok[t]
// set (inverse interpolation of the secular function around a bracket).
__device__ __forceinline__ double inv_interp4(const double* X, const double* Y, const bool* ok) {
  double c = 0.0;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    double w = 1.0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q != p && ok[q]) w *= Y[q] / (Y[q] - Y[p]);
    c += ok[p] ? w * X[p] : 0.0;
  }
  return c;
}

// Lanes of a group that check the window around the interpolated root in a
// verification round (cstar -/+ 0.49 tol: the two points bracket the root in
// less than tol when the estimate is within 0.49 tol of it, which the inverse
// cubic through grid points 0.005 km/s apart reaches with orders of magnitude
// to spare; a miss is retried once around a Newton step); the other lanes
// scan the next period.
constexpr int DK_NW = 2;

// (at most 256 VGPRs: two waves per SIMD on batches with enough columns)
template <int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_forward(BatchDev B, int col_begin, int col_end, int persistent, int lds_stride) {
  extern __shared__ double4 lds[];
  const int lane = threadIdx.x & 63;
  const int r = lane & (S - 1);
  const int gbase = lane & ~(S - 1);
  const int gpb = blockDim.x / S;  // groups per block
  const int gid_blk = threadIdx.x / S;
  double4* mdl = lds + (size_t)gid_blk * lds_stride;
  const double dc = (double)0.005f;
  const double onea = 1.5;
  const double twopi = 2.0 * 3.141592653589793;
  const double tol = B.tol;

  // w: the group's position in the scheduling order, col: its column.  The
  // first column of every group is its group index; a persistent group then
  // takes positions ngroups, ngroups + 1, ... from the work counter (no burst
  // of same-address atomics when the launch starts).
  const int ngroups = gridDim.x * gpb;
  int w = col_begin + blockIdx.x * gpb + gid_blk, col = 0;

  // group state (identical in all S lanes of the group)
  int L = 0, K = 0, k = 0, st = ST_DONE;
  int64_t loff = 0, toff = 0;
  double cm = 0, c1 = 0, clow = 0, omega = 1, betmx_d = 0;
  FV d1 = {0, 0};
  double d1x = 0;  // abscissa of the stored del1 value
  int s1st = 0;    // sign bit of del1st
  double lo = 0, hi = 0;
  FV flo = {0, 0}, fhi = {0, 0};
  int slo = 0, lo_ok = 0, hi_ok = 0;
  int nref = 0;     // multisection rounds of the current bracket
  int nround = 0;   // evaluation rounds of the current column (hang guard)
  double cstar = 0, c1n = 0, omega_n = 1;  // verification round: estimate, next scan start
  double pcur = 1, pnext = 1;              // periods k and k+1
  int nver = 0;                            // verification rounds of the current period

  auto begin_column = [&]() {
    if (w >= col_end) {
      st = ST_DONE;
      return;
    }
    col = B.order ? B.order[w] : w;
    col_geom(B, col, &L, &loff, &K, &toff);
    for (int m = r; m < L; m += S)
      layer_rec(B.layers[loff + m], B.layers[loff + max(m - 1, 0)], m == 0, &mdl[2 * m],
                &mdl[2 * m + 1]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (B.bad[col] != 0 || K <= 0) {
      // bad model: every period is zero-filled
      for (int kk = r; kk < K; kk += S) {
        int64_t s = solve_index(B, col, toff, kk);
        B.croot[s] = 0.0;
        B.pv[s] = 0.0;
      }
      if (r == 0)
        B.status[col] = B.bad[col] == DK_PREP_GEOMETRY ? DK_STATUS_GEOMETRY
                        : B.bad[col]                    ? DK_STATUS_BADMODEL
                                                        : 0;
      st = ST_DONE;
      return;
    }
    cm = B.cc[col];
    betmx_d = (double)B.betmx[col];
    nround = 0;
    k = 0;
    c1 = cm;
    clow = cm;
    pcur = period_of(B, toff, 0);
    omega = twopi / pcur;
    pnext = K > 1 ? period_of(B, toff, 1) : 1.0;
    st = ST_START;
  };
  begin_column();

  auto fail_from = [&](int kf) {
    for (int kk = kf + r; kk < K; kk += S) {
      int64_t s = solve_index(B, col, toff, kk);
      B.croot[s] = 0.0;
      B.pv[s] = 0.0;
    }
    if (r == 0) B.status[col] = kf + 1;
    st = ST_DONE;
  };

  // stores the root of period k and moves to period k+1; false when the
  // column is finished (all periods done, or the root lies above betmx)
  auto emit = [&](double root) -> bool {
    if (root > betmx_d) {
      fail_from(k);
      return false;
    }
    if (r == 0) {
      int64_t s = solve_index(B, col, toff, k);
      B.croot[s] = root;
      B.pv[s] = (double)(float)root;
    }
    k++;
    if (k >= K) {
      if (r == 0) B.status[col] = 0;
      st = ST_DONE;
      return false;
    }
    // period k is in pnext; period k+1 is read now, a round or more before
    // its first use
    pcur = pnext;
    pnext = k + 1 < K ? period_of(B, toff, k + 1) : 1.0;
    return true;
  };

  while (true) {
    // a wave runs while any of its groups has work; finished groups in a
    // persistent launch fetch the next column
    if (persistent && st == ST_DONE && w < col_end) {
      int c0 = 0;
      if (r == 0) c0 = atomicAdd(B.work, 1);
      w = __shfl(c0, gbase, 64) + col_begin + ngroups;
      begin_column();
    }
    if (!__any(st != ST_DONE)) break;

    // ---- trial points: lane r evaluates trial(r); the search logic below
    // recomputes any other lane's point from the same expression and the
    // state at the round's start (the same bits, and no cross-lane read).
    // Every state's grid is xb + (j + jo) * xs (a decreasing scan as a
    // negative step: c1 - a*dc and c1 + a*(-dc) round alike); the
    // verification window's four points are cstar + off * tol.
    double xb = 1.0, xs = 0.0;
    int jo = 0;
    if (st == ST_START) {
      xb = c1;
      xs = dc;
    } else if (st == ST_UP) {
      xb = c1;
      xs = dc;
      jo = 1;
    } else if (st == ST_DN) {
      xb = c1;
      xs = -dc;
      jo = 1;
    } else if (st == ST_REF) {
      xb = lo;
      xs = (hi - lo) / (double)(S + 1);
      jo = 1;
    } else if (st == ST_VER && k + 1 < K) {
      xb = c1n;
      xs = dc;
      jo = -DK_NW;
    }
    const bool ver_r = st == ST_VER;
    const double cstar_r = cstar;
    auto trial = [&](int j) -> double {
      if (ver_r && j < DK_NW) {
        return cstar_r + (j == 0 ? -0.49 : 0.49) * tol;
      }
      return xb + (double)(j + jo) * xs;
    };
    const double x = trial(r);
    double om = omega;
    bool elig = true;
    if (st == ST_DN) {
      elig = x > clow;
    } else if (st == ST_VER) {
      if (r < DK_NW)
        elig = (x > lo) && (x < hi);
      else if (k + 1 < K)
        om = omega_n;
      else
        elig = false;
    } else if (st != ST_START && st != ST_UP && st != ST_REF) {
      elig = false;
    }
    FV f = {1.0, 0};
    if (st != ST_DONE && elig) f.m = secular_fwd(x, om, L, mdl, &f.e);
    const int sg = signbit(f.m) ? 1 : 0;
    const uint64_t neg = group_bits<S>(__ballot(sg), gbase);
    const uint64_t okm = group_bits<S>(__ballot(elig), gbase);
    const bool fl = (x < cm) || (x >= betmx_d + dc);
    const uint64_t flall = group_bits<S>(__ballot(fl), gbase);
    auto sign_at = [&](int j) { return (int)((neg >> j) & 1ull); };
    auto fv_at = [&](int j) { return FV{__shfl(f.m, gbase + j, 64), __shfl(f.e, gbase + j, 64)}; };
    auto x_at = [&](int j) { return trial(j); };
    // sign changes between consecutive trial points j-1, j of the first n
    // (p0: the sign before point 0)
    auto changes = [&](uint64_t s, int p0, int n) {
      const uint64_t m = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
      s &= m;
      return (s ^ ((s << 1) | (uint64_t)p0)) & m;
    };

    if (st == ST_DONE) continue;
    // every state is bounded by the betmx test, the clow floor or the bracket
    // width; this cap only guards against non-finite arithmetic
    if (++nround > DK_MAX_ROUNDS) {
      fail_from(k);
      continue;
    }
    // the lanes' trial points belong to the state this round started in
    const int st0 = st;

    // Upward scan over the n lanes from `base` (the reference's getsol steps,
    // surfdisp96.f:384-476); first = 1 when lane `base` is c1 itself (the
    // period's first step, which also sets the scan direction).  A bracket
    // leads to a verification round around the interpolated root.
    auto scan = [&](int base, int n, int first) {
      if (first) {
        int s0 = sign_at(base);
        d1 = fv_at(base);
        d1x = c1;
        int idir = +1;
        if (k == 0)
          s1st = s0;
        else if (s0 != s1st)
          idir = -1;
        if (idir < 0) {
          st = ST_DN;
          return;
        }
      }
      // changes over j = first .. n-1 (j = 0 against the stored del1 sign)
      uint64_t chg = changes(neg >> base, signbit(d1.m) ? 1 : 0, n);
      if (first) chg &= ~1ull;
      uint64_t flm = (flall >> base) & ((n >= 64) ? ~0ull : ((1ull << n) - 1ull));
      if (first) flm &= ~1ull;
      int jc = ffs64(chg), jf = ffs64(flm);
      if (jc >= 0 && (jf < 0 || jc <= jf)) {
        lo = (jc == 0) ? c1 : x_at(base + jc - 1);
        flo = (jc == 0) ? d1 : fv_at(base + jc - 1);
        slo = (jc == 0) ? (signbit(d1.m) ? 1 : 0) : sign_at(base + jc - 1);
        lo_ok = (jc == 0) ? (d1x == c1) : 1;
        hi = x_at(base + jc);
        fhi = fv_at(base + jc);
        hi_ok = 1;
        nref = 0;
        if (!lo_ok) {
          st = ST_REF;
          return;
        }
        // inverse interpolation through the grid points jc-2 .. jc+1 (j = -1
        // is the stored del1 point of an upward continuation)
        double PX[4], PY[4];
        int PE[4];
        bool ok[4];
        int emax = -2147483647;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int j = jc - 2 + t;
          const int jj = min(max(j, 0), n - 1);
          PX[t] = x_at(base + jj);
          FV fj = fv_at(base + jj);
          ok[t] = (j >= 0 && j < n);
          if (j == -1 && !first && d1x == c1) {
            PX[t] = c1;
            fj = d1;
            ok[t] = true;
          }
          PY[t] = fj.m;
          PE[t] = fj.e;
          if (ok[t]) emax = max(emax, fj.e);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) PY[t] = ok[t] ? ldexp(PY[t], PE[t] - emax) : 0.0;
        double cs = inv_interp4(PX, PY, ok);
        if (!(cs > lo && cs < hi)) cs = fv_ratio_root(lo, hi, flo, fhi);
        cstar = cs;
        c1n = cs - onea * dc;
        omega_n = (k + 1 < K) ? twopi / pnext : omega;
        nver = 0;
        st = ST_VER;
      } else if (jf >= 0) {
        fail_from(k);
      } else {
        c1 = x_at(base + n - 1);
        d1 = fv_at(base + n - 1);
        d1x = c1;
        st = ST_UP;
      }
    };

    // scan lanes of this round: all S lanes, or the lanes after the window of
    // a verification round that accepted its root
    int scan_base = -1, scan_first = 0;
    if (st == ST_START || st == ST_UP) {
      scan_base = 0;
      scan_first = (st == ST_START);
    } else if (st == ST_VER) {
      // the window points inside (lo, hi), ascending: first sign change
      double vlo = lo, vhi = hi;
      FV vflo = flo, vfhi = fhi;
      int vslo = slo;
      bool found = false;
      for (int j = 0; j < DK_NW; j++) {
        double xj = x_at(j);
        FV fj = fv_at(j);
        if (!found && ((okm >> j) & 1ull)) {
          if (sign_at(j) != vslo) {
            vhi = xj;
            vfhi = fj;
            found = true;
          } else {
            vlo = xj;
            vflo = fj;
          }
        }
      }
      lo = vlo;
      hi = vhi;
      flo = vflo;
      fhi = vfhi;
      nver++;
      // a missed window: one Newton step from the window point next to the
      // root with the window's own slope, verified once more
      double cnew = -1.0;
      if (hi - lo > tol && nver == 1) {
        const int j0 = found ? 0 : DK_NW - 2;  // root below / above the window
        if (((okm >> j0) & 3ull) == 3ull) {
          const double xa = x_at(j0), xb = x_at(j0 + 1);
          const FV fa = fv_at(j0), fb = fv_at(j0 + 1);
          const int em = max(fa.e, fb.e);
          const double va = ldexp(fa.m, fa.e - em), vb = ldexp(fb.m, fb.e - em);
          const double slope = (vb - va) / (xb - xa);
          const double xn = found ? xa : xb, vn = found ? va : vb;
          const double cn = xn - vn / slope;
          if (cn > lo && cn < hi) cnew = cn;
        }
      }
      if (cnew > 0.0) {
        cstar = cnew;
        c1n = cnew - onea * dc;
      } else if (hi - lo <= tol) {
        double root = (signbit(flo.m) != signbit(fhi.m)) ? fv_ratio_root(lo, hi, flo, fhi)
                                                          : 0.5 * (lo + hi);
        if (emit(root)) {
          // the next period's first scan step ran in the same round
          omega = omega_n;
          c1 = c1n;
          clow = cm;
          scan_base = DK_NW;
          scan_first = 1;
        }
      } else {
        nref = 0;
        st = ST_REF;
      }
    }
    if (scan_base >= 0) {
      scan(scan_base, S - scan_base, scan_first);
      continue;
    }
    if (st0 == ST_VER || st == ST_DONE) continue;
    if (st == ST_DN) {
      int nv = __popcll(okm);  // eligible lanes form a prefix
      const uint64_t chg = changes(neg, signbit(d1.m) ? 1 : 0, nv);
      int jc = ffs64(chg);
      if (jc >= 0) {
        // bracket (x_jc, x_jc-1): lower end is the new point
        double xb = (jc == 0) ? c1 : x_at(jc - 1);
        FV fb = (jc == 0) ? d1 : fv_at(jc - 1);
        int bok = (jc == 0) ? (d1x == c1) : 1;
        lo = x_at(jc);
        flo = fv_at(jc);
        slo = sign_at(jc);
        lo_ok = 1;
        hi = xb;
        fhi = fb;
        hi_ok = bok;
        nref = 0;
        st = ST_REF;
      } else if (nv < S) {
        // floor reached: continue upward from clow with the last value kept
        if (nv > 0) {
          d1 = fv_at(nv - 1);
          d1x = x_at(nv - 1);
        }
        c1 = clow;
        st = ST_UP;
        continue;
      } else {
        c1 = x_at(S - 1);
        d1 = fv_at(S - 1);
        d1x = c1;
        if (c1 < cm) fail_from(k);
        continue;
      }
    } else if (st == ST_REF) {
      const uint64_t chg = changes(neg, slo, S);
      int jc = ffs64(chg);
      if (jc >= 0) {
        if (jc > 0) {
          lo = x_at(jc - 1);
          flo = fv_at(jc - 1);
          slo = sign_at(jc - 1);
          lo_ok = 1;
        }
        hi = x_at(jc);
        fhi = fv_at(jc);
        hi_ok = 1;
      } else {
        lo = x_at(S - 1);
        flo = fv_at(S - 1);
        slo = sign_at(S - 1);
        lo_ok = 1;
      }
    }

    if (st == ST_REF) nref++;
    if (st == ST_REF && ((hi - lo) <= tol || nref >= DK_MAX_REFINE)) {
      double root;
      bool opp = (signbit(flo.m) != signbit(fhi.m));
      if (lo_ok && hi_ok && opp)
        root = fv_ratio_root(lo, hi, flo, fhi);
      else
        root = 0.5 * (lo + hi);
      if (!emit(root)) continue;
      omega = twopi / pcur;
      c1 = root - onea * dc;
      clow = cm;
      st = ST_START;
    }
  }
}

// ------------------------------------------------------------------ kernel pass
// Implicit differentiation at the root, dc/dm = -(dF/dm)/(dF/dc), on the
// potential-basis recursion of dk_fdm.h (the value is dltar4's times a
// positive factor, which drops out at F = 0).
//   sweep 1, bottom-up: the state and its tangent in c; the state entering
//            the bottom layer of every segment of KERN_SEG layers (counted
//            from the top) but the lowest is kept as a checkpoint with its
//            exponent (the lowest segment starts from the half-space, which
//            sweep 2 evaluates again);
//   sweep 2, top-down, a segment at a time: the segment's states are
//            recomputed from its checkpoint (value steps only) into LDS, then
//            the adjoint dF/d(state) runs through it.  At layer m the state
//            is propagated through the layer again (its wave functions are
//            needed for the adjoint anyway), the interface above it and the
//            layer are differentiated, and the partials in the flattened
//            (a, b, rho) are chained to the unflattened layer.  The interface
//            above layer m also depends on layer m-1's (rho, b), so each
//            layer's partials are complete, and emitted, one step later.
// Checkpoints and outputs are written and read with non-temporal accesses
// (streamed past the caches, which keep the layer records instead).
// Checkpoint layout for segment q: state entry i at tape[(q*5 + i) * nthr + tid],
// its exponent at ((int*)(tape + nq*5*nthr))[q * nthr + tid] (nq checkpoints
// per thread).  LDS segment buffer of the thread (blockDim.x = bd
// threads): state entry (j, i) at seg[(j*5 + i) * bd], exponent j at
// sege[j * bd].  States are stored with an exponent of their own, so the
// recomputation may normalise them on another schedule than sweep 1: the
// represented values are the same (exact powers of two).
__device__ __forceinline__ void kern_solve(const BatchDev& B, const KernDev& O, int col, int k,
                                           int64_t s, double* __restrict__ tape, int64_t nthr,
                                           int tid, double* __restrict__ seg,
                                           int* __restrict__ sege, int bd, int nq) {
  int* __restrict__ tape_e = (int*)(tape + (int64_t)nq * 5 * nthr);
  int L;
  int64_t loff, toff;
  int K;
  col_geom(B, col, &L, &loff, &K, &toff);
  const double c = B.croot[s];
  const int nout = O.node_mode ? O.nz : L;
  auto out_index = [&](int m) -> int64_t {
    if (O.node_mode) return (int64_t)col + (int64_t)B.ncol * ((int64_t)k + (int64_t)K * m);
    return s + O.nsolves * (int64_t)m;
  };
  if (!(c > 0.0)) {
    for (int m = 0; m < nout; m++) {
      int64_t o = out_index(m);
      if (O.dvs) __builtin_nontemporal_store(0.0, O.dvs + o);
      if (O.dvp) __builtin_nontemporal_store(0.0, O.dvp + o);
      if (O.drho) __builtin_nontemporal_store(0.0, O.drho + o);
    }
    return;
  }
  const double4* lay = B.layers + loff;
  const double* btp = B.btp + loff;
  const double* btr = B.btr + loff;
  double omega = 2.0 * 3.141592653589793 / period_of(B, toff, k);
  if (omega < 1.0e-4) omega = 1.0e-4;
  const double wk = omega / c, c2 = c * c, ic2 = 1.0 / c2, dwk = -wk / c;
  const int H = L - 1;
  auto cl = [&](int i) { return min(max(i, 0), H); };  // clamped record index

  // ---- sweep 1
  double v[5], dv[5];
  {
    const double4 lh = lay[H];
    double ran, rbn, sgA, sgB;
    fdm_start(c2, lh, v, &ran, &rbn, &sgA, &sgB);
    const double dra = start_dr(ran, sgA, 2.0 * c / (lh.y * lh.y));
    const double drb = start_dr(rbn, sgB, 2.0 * c / (lh.z * lh.z));
    dv[0] = fma(dra, rbn, ran * drb);
    dv[1] = 0.0;
    dv[2] = 0.0;
    dv[3] = -dra;
    dv[4] = drb;
  }
  auto iface_t = [&](const double4& up, const double4& lo) {
    double eps, dmu, yz[4], w[5], dw[5];
    iface_consts(up, lo, &eps, &dmu);
    const double eta = dmu * ic2;
    fdm_iface_fn(eps, eta, v, yz, w);
    fdm_iface_tan(eps, eta, -2.0 * eta / c, v, dv, yz, dw);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      v[i] = w[i];
      dv[i] = dw[i];
    }
  };
  // layer records one iteration ahead (loads before the iteration's tape
  // stores: a load issued after them would wait for them too)
  double4 lcur = lay[cl(H - 1)];
  if (H >= 1) iface_t(lcur, lay[H]);
  double4 lup = lay[cl(H - 2)];
  int Ef = 0;
  for (int m = H - 1; m >= 0; m--) {
    const double4 l = lcur, lu = lup;
    lcur = lup;
    lup = lay[cl(m - 2)];
    if (m % KERN_SEG == KERN_SEG - 1 && m != H - 1) {  // the bottom of segment m / KERN_SEG
      const int q = m / KERN_SEG;
      double* tp = tape + (int64_t)(q * 5) * nthr + tid;
#pragma unroll
      for (int i = 0; i < 5; i++) __builtin_nontemporal_store(v[i], tp + (int64_t)i * nthr);
      __builtin_nontemporal_store(Ef, tape_e + (int64_t)q * nthr + tid);
    }
    const double kd = wk * l.x, dkd = dwk * l.x;
    WaveFn A, Bw;
    wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
    wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, Bw);
    double dA[4], dB[4], t[4], u[5], du[5];
    wave_tan(A, kd, -2.0 * c / (l.y * l.y), dkd, dA);
    wave_tan(Bw, kd, -2.0 * c / (l.z * l.z), dkd, dB);
    fdm_layer_fn(A, Bw, v, t, u);
    fdm_layer_tan(A, Bw, dA, dB, v, dv, t, du);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      v[i] = u[i];
      dv[i] = du[i];
    }
    if (m >= 1) iface_t(lu, l);
    if ((m & 1) == 0 || m == H - 1) {
      // values and tangent by the same exact power of two
      double mx = 0.0;
#pragma unroll
      for (int i = 0; i < 5; i++) mx = fmax(mx, fabs(v[i]));
      int kk = 0;
      if (mx > 1e-300 && isfinite(mx)) frexp(mx, &kk);
#pragma unroll
      for (int i = 0; i < 5; i++) {
        v[i] = ldexp(v[i], -kk);
        dv[i] = ldexp(dv[i], -kk);
      }
      Ef += kk;
    }
  }
  const int EF = Ef;  // exponent of the surface state: reference scale
  const double4 l0 = lay[0];
  const double ib0 = 1.0 / (l0.z * l0.z);
  const double tt = fma(-c2, ib0, 2.0);
  const double tb = -fma(2.0 * tt, v[2], 4.0 * v[1]);  // dF/dt
  const double dFdc = -fma(tt, fma(tt, dv[2], 4.0 * dv[1]), 4.0 * dv[0]) + tb * (-2.0 * c * ib0);
  const double sc = (dFdc != 0.0) ? -1.0 / dFdc : 0.0;

  // ---- sweep 2
  double g[5] = {-4.0, -4.0 * tt, -tt * tt, 0.0, 0.0};
  int Eg = 0;
  // pending partials (flattened a, b, rho) of the previous layer, in kernel
  // units; the surface's t depends on b_0
  double pa = 0.0, pb = tb * 2.0 * c2 * ib0 / l0.z * sc, pr = 0.0;
  int node = 0, sub = 0;
  double acc_lo[3] = {0, 0, 0}, acc_hi[3] = {0, 0, 0};
  int ns_cur = (O.node_mode && O.nz > 1) ? O.nsub[0] : 1;
  // emission of layer m's partials: the flattening chain (a_flat = a btp,
  // b_flat = b btp, rho_flat = rho btp^-2.275), then per layer or per node
  auto emit = [&](int m, double qa, double qb, double qr, double bt, double brt) {
    const double ea = qa * bt, eb = qb * bt, er = qr * brt;
    if (!O.node_mode) {
      int64_t o = out_index(m);
      if (O.dvp) __builtin_nontemporal_store(ea, O.dvp + o);
      if (O.dvs) __builtin_nontemporal_store(eb, O.dvs + o);
      if (O.drho) __builtin_nontemporal_store(er, O.drho + o);
      return;
    }
    // node mode: sublayer j (1-based) of interval `node` with ns sublayers
    if (m < H) {
      const int ns = ns_cur;
      int j = sub + 1;
      double whi = (double)(2 * j - 1) / (double)(2 * ns);
      double wlo = 1.0 - whi;
      acc_lo[0] += wlo * eb;
      acc_lo[1] += wlo * ea;
      acc_lo[2] += wlo * er;
      acc_hi[0] += whi * eb;
      acc_hi[1] += whi * ea;
      acc_hi[2] += whi * er;
      sub++;
      if (sub == ns && node < O.nz - 1) {
        int64_t o = out_index(node);
        if (O.dvs) __builtin_nontemporal_store(acc_lo[0], O.dvs + o);
        if (O.dvp) __builtin_nontemporal_store(acc_lo[1], O.dvp + o);
        if (O.drho) __builtin_nontemporal_store(acc_lo[2], O.drho + o);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          acc_lo[q] = acc_hi[q];
          acc_hi[q] = 0.0;
        }
        node++;
        sub = 0;
        ns_cur = (node < O.nz - 1) ? O.nsub[node] : 1;
      }
    } else {
      int64_t o = out_index(node);
      if (O.dvs) __builtin_nontemporal_store(acc_lo[0] + eb, O.dvs + o);
      if (O.dvp) __builtin_nontemporal_store(acc_lo[1] + ea, O.dvp + o);
      if (O.drho) __builtin_nontemporal_store(acc_lo[2] + er, O.drho + o);
    }
  };
  // records and flattening factors one iteration ahead (unconditional loads)
  double4 ln = l0;
  double4 lprev = make_double4(0.0, 0.0, l0.z, l0.w);  // (rho, b) of the layer above
  double btn = btp[0], brn = btr[0], bt_prev = 0.0, br_prev = 0.0;
  for (int m0 = 0; m0 < H; m0 += KERN_SEG) {
    const int m1 = min(m0 + KERN_SEG - 1, H - 1);
    {
      // the segment's states: its checkpoint (entering layer m1), then up;
      // the bottom segment's starts from the half-space as sweep 1 does
      double v[5];
      int Ev = 0;
      if (m1 == H - 1) {
        const double4 lh = lay[H];
        double ran, rbn, sgA, sgB, eps, dmu, yz[4], w[5];
        fdm_start(c2, lh, v, &ran, &rbn, &sgA, &sgB);
        iface_consts(lay[H - 1], lh, &eps, &dmu);
        fdm_iface_fn(eps, dmu * ic2, v, yz, w);
#pragma unroll
        for (int i = 0; i < 5; i++) v[i] = w[i];
      } else {
        const int q = m0 / KERN_SEG;
        const double* tp = tape + (int64_t)(q * 5) * nthr + tid;
#pragma unroll
        for (int i = 0; i < 5; i++) v[i] = __builtin_nontemporal_load(tp + (int64_t)i * nthr);
        Ev = __builtin_nontemporal_load(tape_e + (int64_t)q * nthr + tid);
      }
      auto put = [&](int j) {
#pragma unroll
        for (int i = 0; i < 5; i++) seg[(j * 5 + i) * bd] = v[i];
        sege[j * bd] = Ev;
      };
      put(m1 - m0);
      double4 l = lay[m1];
      for (int m = m1; m > m0; m--) {
        const double4 lu = lay[m - 1];
        const double kd = wk * l.x;
        WaveFn A, Bw;
        wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
        wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, Bw);
        double t[4], u[5], eps, dmu, yz[4];
        fdm_layer_fn(A, Bw, v, t, u);
        iface_consts(lu, l, &eps, &dmu);
        fdm_iface_fn(eps, dmu * ic2, u, yz, v);
        Ev += pow2_normalize(v);
        put(m - 1 - m0);
        l = lu;
      }
    }
    for (int m = m0; m <= m1; m++) {
    const double4 l = ln;
    double tc[6];
#pragma unroll
    for (int i = 0; i < 5; i++) tc[i] = seg[((m - m0) * 5 + i) * bd];
    tc[5] = (double)sege[(m - m0) * bd];
    const double bt = btn, brt = brn;
    ln = lay[m + 1];
    btn = btp[m + 1];
    brn = btr[m + 1];
    const double kd = wk * l.x;
    WaveFn A, Bw;
    wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
    wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, Bw);
    double t[4], u[5], ub[5], vb[5], ab[4], bb[4];
    fdm_layer_fn(A, Bw, tc, t, u);
    const double scl = ldexp(sc, (int)tc[5] + Eg - EF);
    double ca = 0.0, cb = 0.0, cr = 0.0;  // this layer's partials
    if (m >= 1) {
      double eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
      iface_consts(lprev, l, &eps, &dmu);
      fdm_iface_fn(eps, dmu * ic2, u, yz, w);
      fdm_iface_adj(eps, dmu * ic2, u, yz, g, ub, &epsb, &etab);
      iface_chain(lprev, l, eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
      emit(m - 1, pa, fma(bu, scl, pb), fma(ru, scl, pr), bt_prev, br_prev);
      cb = bl;
      cr = rl;
    } else {
#pragma unroll
      for (int i = 0; i < 5; i++) ub[i] = g[i];
    }
    fdm_layer_adj(A, Bw, tc, t, ub, vb, ab, bb);
    double Xa, kda, Xb, kdb;
    wave_adj(A, kd, ab, &Xa, &kda);
    wave_adj(Bw, kd, bb, &Xb, &kdb);
    ca = Xa * 2.0 * c2 / (l.y * l.y * l.y);
    cb += Xb * 2.0 * c2 / (l.z * l.z * l.z);
    if (m == 0) {
      pa = ca * scl;
      pb = fma(cb, scl, pb);
      pr = cr * scl;
    } else {
      pa = ca * scl;
      pb = cb * scl;
      pr = cr * scl;
    }
    bt_prev = bt;
    br_prev = brt;
    lprev = make_double4(0.0, 0.0, l.z, l.w);
#pragma unroll
    for (int i = 0; i < 5; i++) g[i] = vb[i];
    if (m & 1) Eg += pow2_normalize(g);
    }
  }
  // the half-space: the interface above it, then its start state
  double ub[5];
  double ha = 0.0, hb = 0.0, hr = 0.0;
  const double sch = ldexp(sc, Eg - EF);
  const double4 lh = lay[H];
  double v0[5], ran, rbn, sgA, sgB;
  fdm_start(c2, lh, v0, &ran, &rbn, &sgA, &sgB);
  if (H >= 1) {
    double eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
    iface_consts(lprev, lh, &eps, &dmu);
    fdm_iface_fn(eps, dmu * ic2, v0, yz, w);
    fdm_iface_adj(eps, dmu * ic2, v0, yz, g, ub, &epsb, &etab);
    iface_chain(lprev, lh, eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
    emit(H - 1, pa, fma(bu, sch, pb), fma(ru, sch, pr), bt_prev, br_prev);
    hb = bl * sch;
    hr = rl * sch;
  } else {
#pragma unroll
    for (int i = 0; i < 5; i++) ub[i] = g[i];
    hb = pb;  // the surface term of layer 0, the half-space itself
  }
  const double rab = fma(ub[0], rbn, -ub[3]), rbb = fma(ub[0], ran, ub[4]);
  ha = rab * start_dr(ran, sgA, -2.0 * c2 / (lh.y * lh.y * lh.y)) * sch;
  hb = fma(rbb * start_dr(rbn, sgB, -2.0 * c2 / (lh.z * lh.z * lh.z)), sch, hb);
  emit(H, ha, hb, hr, btn, brn);
}

__global__ void __launch_bounds__(KERN_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_kernels(BatchDev B, KernDev O, double* tape, int nq) {
  extern __shared__ double kseg_lds[];
  const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int bd = blockDim.x;
  double* seg = kseg_lds + threadIdx.x;
  int* sege = (int*)(kseg_lds + KERN_SEG * 5 * bd) + threadIdx.x;
  for (int64_t t = tid; t < O.nsolves; t += nthr) {
    const int64_t s = O.solve_order ? (int64_t)O.solve_order[t] : t;
    if (s < 0 || s >= O.nsolves) continue;  // (an order of inconsistent nk, toff)
    int col, k;
    if (B.uniform) {
      col = (int)(s % B.ncol);
      k = (int)(s / B.ncol);
    } else {
      col = O.solve_col[s];
      k = (int)(s - B.toff[col]);
    }
    kern_solve(B, O, col, k, s, tape, nthr, tid, seg, sege, bd, nq);
  }
}

// ------------------------------------------------------------------ VJP kernel
// grad_m = sum_k grad_c[s(col,k)] dc_k/dm for the unflattened layer vs, vp, rho
// of a ragged batch (outputs laid out like the model, at loff[col] + m).
// Two columns per wavefront, one per half (VJP_LANES lanes), lane k of a half
// owning period k (columns with more than VJP_LANES periods take several
// chunks).  Sweep 1 carries the tangent dF/dc with the tape; sweep 2 is the
// adjoint with (alpha, beta, rho) duals, already scaled by -grad_c/(dF/dc);
// per-layer sums reduce across the half's lanes with shuffles, so the result
// does not depend on scheduling or on the other columns of the batch.
constexpr int VJP_LANES = 32;
__device__ __forceinline__ double half_sum(double v) {
#pragma unroll
  for (int o = VJP_LANES / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(64) k_vjp(BatchDev B, const double* __restrict__ grad_c,
                                            double* __restrict__ gvs, double* __restrict__ gvp,
                                            double* __restrict__ grho, double* __restrict__ gthk,
                                            const float* __restrict__ thk32,
                                            const double* __restrict__ thk64,
                                            double* __restrict__ tape) {
  const int lane = threadIdx.x & (VJP_LANES - 1);
  const int half = threadIdx.x / VJP_LANES;
  constexpr int CPW = 64 / VJP_LANES;  // columns per wavefront
  const long nslot = (long)gridDim.x * 64;
  const long slot = (long)blockIdx.x * 64 + threadIdx.x;
  for (int col = blockIdx.x * CPW + half; col < B.ncol; col += gridDim.x * CPW) {
    int L, K;
    int64_t loff, toff;
    col_geom(B, col, &L, &loff, &K, &toff);
    const double4* lay = B.layers + loff;
    const double* btp = B.btp + loff;
    const double* btr = B.btr + loff;
    // a column without periods still writes its (zero) gradients
    for (int k0 = 0; k0 == 0 || k0 < K; k0 += VJP_LANES) {
      const int k = k0 + lane;
      double c = 0.0, gc = 0.0;
      if (k < K) {
        int64_t s = solve_index(B, col, toff, k);
        c = B.croot[s];
        gc = grad_c[s];
      }
      const bool act = (c > 0.0) && (gc != 0.0) && isfinite(gc);
      double omega = 1.0;
      if (act) {
        omega = 2.0 * 3.141592653589793 / period_of(B, toff, k);
        if (omega < 1.0e-4) omega = 1.0e-4;
      }
      // ---- sweep 1: the state (taped) and its tangent in c, as in the
      // kernel pass (dk_fdm.h)
      const int H = L - 1;
      auto cl = [&](int i) { return min(max(i, 0), H); };
      // (lanes without a solve compute nothing below: c = 0 there)
      const double wk = act ? omega / c : 0.0, c2 = c * c, ic2 = act ? 1.0 / c2 : 1.0,
                   dwk = act ? -wk / c : 0.0;
      double lane_scale = 0.0, tb = 0.0, ib0 = 0.0;
      int EF = 0;
      if (act) {
        double v[5], dv[5];
        {
          const double4 lh = lay[H];
          double ran, rbn, sgA, sgB;
          fdm_start(c2, lh, v, &ran, &rbn, &sgA, &sgB);
          const double dra = start_dr(ran, sgA, 2.0 * c / (lh.y * lh.y));
          const double drb = start_dr(rbn, sgB, 2.0 * c / (lh.z * lh.z));
          dv[0] = fma(dra, rbn, ran * drb);
          dv[1] = 0.0;
          dv[2] = 0.0;
          dv[3] = -dra;
          dv[4] = drb;
        }
        auto iface_t = [&](const double4& up, const double4& lo) {
          double eps, dmu, yz[4], w[5], dw[5];
          iface_consts(up, lo, &eps, &dmu);
          const double eta = dmu * ic2;
          fdm_iface_fn(eps, eta, v, yz, w);
          fdm_iface_tan(eps, eta, -2.0 * eta / c, v, dv, yz, dw);
#pragma unroll
          for (int i = 0; i < 5; i++) {
            v[i] = w[i];
            dv[i] = dw[i];
          }
        };
        // records one iteration ahead, before the iteration's tape stores
        double4 lcur = lay[cl(H - 1)];
        if (H >= 1) iface_t(lcur, lay[H]);
        double4 lup = lay[cl(H - 2)];
        int Ef = 0;
        for (int m = H - 1; m >= 0; m--) {
          const double4 l = lcur, lu = lup;
          lcur = lup;
          lup = lay[cl(m - 2)];
          double* tp = tape + (long)(m * 6) * nslot + slot;
#pragma unroll
          for (int i = 0; i < 5; i++) tp[(long)i * nslot] = v[i];
          tp[(long)5 * nslot] = (double)Ef;
          const double kd = wk * l.x, dkd = dwk * l.x;
          WaveFn A, Bw;
          wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
          wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, Bw);
          double dA[4], dB[4], t[4], u[5], du[5];
          wave_tan(A, kd, -2.0 * c / (l.y * l.y), dkd, dA);
          wave_tan(Bw, kd, -2.0 * c / (l.z * l.z), dkd, dB);
          fdm_layer_fn(A, Bw, v, t, u);
          fdm_layer_tan(A, Bw, dA, dB, v, dv, t, du);
#pragma unroll
          for (int i = 0; i < 5; i++) {
            v[i] = u[i];
            dv[i] = du[i];
          }
          if (m >= 1) iface_t(lu, l);
          if ((m & 1) == 0 || m == H - 1) {
            double mx = 0.0;
#pragma unroll
            for (int i = 0; i < 5; i++) mx = fmax(mx, fabs(v[i]));
            int kk = 0;
            if (mx > 1e-300 && isfinite(mx)) frexp(mx, &kk);
#pragma unroll
            for (int i = 0; i < 5; i++) {
              v[i] = ldexp(v[i], -kk);
              dv[i] = ldexp(dv[i], -kk);
            }
            Ef += kk;
          }
        }
        EF = Ef;
        const double4 l0 = lay[0];
        ib0 = 1.0 / (l0.z * l0.z);
        const double tt = fma(-c2, ib0, 2.0);
        tb = -fma(2.0 * tt, v[2], 4.0 * v[1]);
        const double dFdc =
            -fma(tt, fma(tt, dv[2], 4.0 * dv[1]), 4.0 * dv[0]) + tb * (-2.0 * c * ib0);
        lane_scale = (dFdc != 0.0) ? -gc / dFdc : 0.0;
      }
      // ---- sweep 2: the adjoint; per-layer partials in the flattened
      // (a, b, rho, d), scaled by -grad_c/(dF/dc), are reduced over the lanes.
      // The thickness chain goes through the layer radii r0 (top) and r1
      // (bottom) of the flattening (d = ar ln(r0/r1), tmp = 2 ar/(r0 + r1)
      // scales a, b and rho^-2.275): R0, R1 = dF/dr0, dF/dr1, and a thickness
      // change moves r1 of its layer and both radii of every deeper layer.
      // Layer m's partials are complete one step later (the interface above
      // layer m+1 depends on them), and are emitted then.
      const double ar = 6370.0;
      auto thk_at = [&](int m) -> double {
        return (m == H) ? 1.0 : (thk64 ? thk64[loff + m] : (double)thk32[loff + m]);
      };
      double prevR1 = 0.0;
      auto emit = [&](int m, const double* part_in, const double4& l, double bt, double btr_m,
                      double r0, double r1) {
        double part[4] = {part_in[0], part_in[1], part_in[2], part_in[3]};
        double R0 = 0.0, R1 = 0.0;
        if (act) {
          const double dtmp = part[0] * (l.y / bt) + part[1] * (l.z / bt) -
                              2.275 * part[2] * (l.w / bt);
          const double ttm = -dtmp * bt / (r0 + r1);
          R0 = ttm + ((m < H) ? part[3] * (ar / r0) : 0.0);
          R1 = ttm - ((m < H) ? part[3] * (ar / r1) : 0.0);
          part[0] *= bt;
          part[1] *= bt;
          part[2] *= btr_m;
        }
        const double sa = half_sum(part[0]), sb = half_sum(part[1]), sr = half_sum(part[2]);
        double s0 = 0.0, s1 = 0.0;
        if (gthk) {
          s0 = half_sum(R0);
          s1 = half_sum(R1);
        }
        if (lane == 0) {
          const int64_t o = loff + m;
          if (k0 == 0) {
            gvp[o] = sa;
            gvs[o] = sb;
            grho[o] = sr;
          } else {
            gvp[o] += sa;
            gvs[o] += sb;
            grho[o] += sr;
          }
          if (gthk) {
            // W[m-1] = R1[m-1] + R0[m]; gthk holds W until the suffix sum below
            if (m > 0) {
              const double w = prevR1 + s0;
              if (k0 == 0)
                gthk[o - 1] = w;
              else
                gthk[o - 1] += w;
            }
            if (m == H) {
              if (k0 == 0)
                gthk[o] = s1;
              else
                gthk[o] += s1;
            }
            prevR1 = s1;
          }
        }
      };
      // lanes without a solve read lane 0's slot (the same cache lines)
      const long tslot = act ? slot : slot - lane;
      auto load_tape = [&](int m, double* t6) {
        const double* tp = tape + (long)(m * 6) * nslot + tslot;
#pragma unroll
        for (int i = 0; i < 6; i++) t6[i] = tp[(long)i * nslot];
      };
      const int mt = H >= 1 ? H - 1 : 0;  // last tape entry
      double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      int Eg = 0;
      double pend[4] = {0.0, 0.0, 0.0, 0.0};  // the previous layer's partials
      if (act) {
        const double tt = fma(-c2, ib0, 2.0);
        g[0] = -4.0;
        g[1] = -4.0 * tt;
        g[2] = -tt * tt;
        pend[1] = tb * 2.0 * c2 * ib0 / lay[0].z * lane_scale;
      }
      double4 ln = lay[0];
      double4 lprev = ln;
      double tn[6];
      load_tape(0, tn);
      double thn = thk_at(0), btn = btp[0], brn = btr[0];
      double bt_prev = 1.0, br_prev = 1.0, r0_prev = ar, r1_prev = ar, Z = 0.0;
      for (int m = 0; m < H; m++) {
        const double4 l = ln;
        double tc[6];
#pragma unroll
        for (int i = 0; i < 6; i++) tc[i] = tn[i];
        const double bt = btn, brt = brn, thk_m = thn;
        ln = lay[m + 1];
        load_tape(min(m + 1, mt), tn);
        btn = btp[m + 1];
        brn = btr[m + 1];
        thn = thk_at(m + 1);
        const double r0 = ar - Z, r1 = r0 - thk_m;
        Z += thk_m;
        double cur[4] = {0.0, 0.0, 0.0, 0.0};
        if (act) {
          const double kd = wk * l.x;
          WaveFn A, Bw;
          wave_fn(fma(-c2, 1.0 / (l.y * l.y), 1.0), kd, A);
          wave_fn(fma(-c2, 1.0 / (l.z * l.z), 1.0), kd, Bw);
          double t[4], u[5], ub[5], vb[5], ab[4], bb[4];
          fdm_layer_fn(A, Bw, tc, t, u);
          const double scl = ldexp(lane_scale, (int)tc[5] + Eg - EF);
          if (m >= 1) {
            double eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
            iface_consts(lprev, l, &eps, &dmu);
            fdm_iface_fn(eps, dmu * ic2, u, yz, w);
            fdm_iface_adj(eps, dmu * ic2, u, yz, g, ub, &epsb, &etab);
            iface_chain(lprev, l, eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
            pend[1] = fma(bu, scl, pend[1]);
            pend[2] = fma(ru, scl, pend[2]);
            cur[1] = bl;
            cur[2] = rl;
          } else {
#pragma unroll
            for (int i = 0; i < 5; i++) ub[i] = g[i];
          }
          fdm_layer_adj(A, Bw, tc, t, ub, vb, ab, bb);
          double Xa, kda, Xb, kdb;
          wave_adj(A, kd, ab, &Xa, &kda);
          wave_adj(Bw, kd, bb, &Xb, &kdb);
          cur[0] = Xa * 2.0 * c2 / (l.y * l.y * l.y) * scl;
          cur[1] = (cur[1] + Xb * 2.0 * c2 / (l.z * l.z * l.z)) * scl;
          cur[2] *= scl;
          cur[3] = (kda + kdb) * wk * scl;
#pragma unroll
          for (int i = 0; i < 5; i++) g[i] = vb[i];
          if (m & 1) Eg += pow2_normalize(g);
        }
        if (m >= 1) emit(m - 1, pend, lprev, bt_prev, br_prev, r0_prev, r1_prev);
        if (m == 0) {
          cur[1] += pend[1];  // the surface term of layer 0
        }
#pragma unroll
        for (int i = 0; i < 4; i++) pend[i] = cur[i];
        lprev = l;
        bt_prev = bt;
        br_prev = brt;
        r0_prev = r0;
        r1_prev = r1;
      }
      // the half-space: the interface above it, then its start state
      {
        const double4 lh = lay[H];
        const double r0 = ar - Z, r1 = r0 - thn;
        double hp[4] = {0.0, 0.0, 0.0, 0.0};
        if (act) {
          double ub[5], v0[5], ran, rbn, sgA, sgB;
          fdm_start(c2, lh, v0, &ran, &rbn, &sgA, &sgB);
          const double sch = ldexp(lane_scale, Eg - EF);
          if (H >= 1) {
            double eps, dmu, yz[4], w[5], epsb, etab, ru, bu, rl, bl;
            iface_consts(lprev, lh, &eps, &dmu);
            fdm_iface_fn(eps, dmu * ic2, v0, yz, w);
            fdm_iface_adj(eps, dmu * ic2, v0, yz, g, ub, &epsb, &etab);
            iface_chain(lprev, lh, eps, dmu, ic2, epsb, etab, &ru, &bu, &rl, &bl);
            pend[1] = fma(bu, sch, pend[1]);
            pend[2] = fma(ru, sch, pend[2]);
            hp[1] = bl * sch;
            hp[2] = rl * sch;
          } else {
#pragma unroll
            for (int i = 0; i < 5; i++) ub[i] = g[i];
            hp[1] = pend[1];  // the surface term: layer 0 is the half-space
          }
          const double rab = fma(ub[0], rbn, -ub[3]), rbb = fma(ub[0], ran, ub[4]);
          hp[0] = rab * start_dr(ran, sgA, -2.0 * c2 / (lh.y * lh.y * lh.y)) * sch;
          hp[1] = fma(rbb * start_dr(rbn, sgB, -2.0 * c2 / (lh.z * lh.z * lh.z)), sch, hp[1]);
        }
        if (H >= 1) emit(H - 1, pend, lprev, bt_prev, br_prev, r0_prev, r1_prev);
        emit(H, hp, lh, btn, brn, r0, r1);
      }
    }
    // dc/dthk[i] = -sum_{j >= i} W[j]; the half-space thickness does not enter
    if (gthk && lane == 0) {
      double run = 0.0;
      for (int m = L - 1; m >= 0; m--) {
        run += gthk[loff + m];
        gthk[loff + m] = -run;
      }
      gthk[loff + L - 1] = 0.0;
    }
  }
}

hipError_t launch_vjp(const BatchDev& B, const double* grad_c, double* gvs, double* gvp,
                      double* grho, double* gthk, const float* thk32, const double* thk64,
                      double* tape, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_vjp, dim3(nblocks), dim3(64), 0, st, B, grad_c, gvs, gvp, grho, gthk, thk32,
                     thk64, tape);
  return hipGetLastError();
}

// ------------------------------------------------------------------ effective dc/dVs
// The G builder's Vs row (CalSurfGAniso_Joint.f90:750-765): with vp and rho
// tied to vs by Brocher, dc/dVs_eff = sen_vs + sen_vp coe_a + sen_rho coe_rho
// (brocher_coeffs, effective_vs).

// One thread per (node, column): eff[(z*K + k)*ncol + col] for z < nz-1 (the
// reference's 1:nz-1 rows) and grad[z*ncol + col] = sum_k gc[k*ncol + col] *
// eff(z, k, col) for every node.  Streams the three node kernels once.
__global__ void k_effective(int ncol, int nz, int K, const float* __restrict__ vel,
                            const double* __restrict__ svs, const double* __restrict__ svp,
                            const double* __restrict__ srho, const double* __restrict__ gc,
                            double* __restrict__ eff, double* __restrict__ grad) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ncol * nz) return;
  const int z = (int)(t / ncol), col = (int)(t % ncol);
  float ca, cr;
  brocher_coeffs(vel[t], &ca, &cr);
  double acc = 0.0;
  for (int k = 0; k < K; k++) {
    const long o = ((long)z * K + k) * ncol + col;
    const double e = effective_vs(svs[o], svp[o], srho[o], ca, cr);
    if (eff && z < nz - 1) eff[o] = e;
    if (grad) acc = fma(gc[(long)k * ncol + col], e, acc);
  }
  if (grad) grad[t] = acc;
}

hipError_t launch_effective(int ncol, int nz, int K, const float* vel, const double* svs,
                            const double* svp, const double* srho, const double* gc, double* eff,
                            double* grad, hipStream_t st) {
  const long n = (long)ncol * nz;
  const int block = 256;
  hipLaunchKernelGGL(k_effective, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, st,
                     ncol, nz, K, vel, svs, svp, srho, gc, eff, grad);
  return hipGetLastError();
}

// ------------------------------------------------------------------ cell sums
// Per-cell sums of a population gradient: acc[(r*nl + l)*ncells + cell[(r*nl +
// l)*ncr + j]] += gl[l*ld + (c - c0)] over the chunk's columns c = r*ncr + j in
// [c0, c0 + n).  One workgroup per (realization, layer) that the chunk touches;
// thread t owns cells t, t + 256, ...  and scans the realization's (cell,
// value) pairs from LDS in column order into eight partial sums, so every sum
// runs in a fixed order (no atomics, the same bits on every run).  Cell ids
// outside [0, ncells) add nowhere.
constexpr int CELL_TILE = 1024;  // a multiple of 8
__global__ void __launch_bounds__(256) k_cell_sum(const double* __restrict__ gl, int64_t ld,
                                                  int nl, int64_t c0, int64_t n, int ncr,
                                                  const int64_t* __restrict__ cell, int ncells,
                                                  double* __restrict__ acc) {
  __shared__ __align__(16) int cid[CELL_TILE];
  __shared__ __align__(16) double val[CELL_TILE];
  const int l = blockIdx.x % nl;
  const int64_t r = c0 / ncr + blockIdx.x / nl;
  const int64_t base = r * ncr;
  const int64_t jlo = max(c0, base) - base, jhi = min(c0 + n, base + ncr) - base;
  const int64_t* crow = cell + (r * nl + l) * (int64_t)ncr;
  const double* grow = gl + (int64_t)l * ld + (base - c0);  // + j: column base + j of the chunk
  for (int cb = 0; cb < ncells; cb += blockDim.x) {
    const int t = cb + threadIdx.x;
    // eight partial sums, column j into sum (j - jlo) mod 8, added in a fixed
    // order at the end (independent chains for the LDS reads to overlap)
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t j0 = jlo; j0 < jhi; j0 += CELL_TILE) {
      const int m = (int)min((int64_t)CELL_TILE, jhi - j0);
      const int mp = (m + 7) & ~7;
      __syncthreads();
      for (int i = threadIdx.x; i < mp; i += blockDim.x) {
        int id = -1;
        double v = 0.0;
        if (i < m) {
          const int64_t c = crow[j0 + i];
          id = (c >= 0 && c < ncells) ? (int)c : -1;
          v = grow[j0 + i];
        }
        cid[i] = id;
        val[i] = v;
      }
      __syncthreads();
      for (int i = 0; i < mp; i += 8) {
        const int4 ca = *reinterpret_cast<const int4*>(&cid[i]);
        const int4 cc = *reinterpret_cast<const int4*>(&cid[i + 4]);
        const double2 v0 = *reinterpret_cast<const double2*>(&val[i]);
        const double2 v1 = *reinterpret_cast<const double2*>(&val[i + 2]);
        const double2 v2 = *reinterpret_cast<const double2*>(&val[i + 4]);
        const double2 v3 = *reinterpret_cast<const double2*>(&val[i + 6]);
        if (ca.x == t) a[0] += v0.x;
        if (ca.y == t) a[1] += v0.y;
        if (ca.z == t) a[2] += v1.x;
        if (ca.w == t) a[3] += v1.y;
        if (cc.x == t) a[4] += v2.x;
        if (cc.y == t) a[5] += v2.y;
        if (cc.z == t) a[6] += v3.x;
        if (cc.w == t) a[7] += v3.y;
      }
    }
    const double sum = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    if (t < ncells) acc[(r * nl + l) * (int64_t)ncells + t] += sum;
  }
}

hipError_t launch_cell_sum(const double* gl, int64_t ld, int nl, int64_t c0, int64_t n, int ncr,
                           const int64_t* cell, int ncells, double* acc, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t nr = (c0 + n - 1) / ncr - c0 / ncr + 1;
  hipLaunchKernelGGL(k_cell_sum, dim3((unsigned)(nr * nl)), dim3(256), 0, st, gl, ld, nl, c0, n,
                     ncr, cell, ncells, acc);
  return hipGetLastError();
}

__global__ void k_zero_int(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ scheduling order
// Counting sort of a ragged batch's columns (and solves) by layer count, then
// period count, largest first: k_order_hist counts columns and solves per
// key, k_order_scan turns the counts into each key's first position,
// k_order_scatter places the items.  work = [column counts | solve counts |
// column bases | solve bases], ORDER_BINS ints each.  The key is
// min(L, ORDER_BINS / 2^kb - 1) * 2^kb + the top kb bits of min(K, 63).
__device__ __forceinline__ int order_key(int L, int K, int kb) {
  const int lk = min(max(L, 0), (ORDER_BINS >> kb) - 1);
  const int kq = kb > 0 ? min(max(K, 0), 63) >> (6 - kb) : 0;
  return (lk << kb) | kq;
}

__global__ void __launch_bounds__(256) k_order_hist(int ncol, const int* Lc, const int* Kc, int kb,
                                                    int* work) {
  __shared__ int hc[ORDER_BINS], hs[ORDER_BINS];
  for (int b = threadIdx.x; b < ORDER_BINS; b += blockDim.x) hc[b] = hs[b] = 0;
  __syncthreads();
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const int K = Kc[c];
    const int key = order_key(Lc[c], K, kb);
    atomicAdd(&hc[key], 1);
    atomicAdd(&hs[key], max(K, 0));
  }
  __syncthreads();
  for (int b = threadIdx.x; b < ORDER_BINS; b += blockDim.x) {
    if (hc[b]) atomicAdd(&work[b], hc[b]);
    if (hs[b]) atomicAdd(&work[ORDER_BINS + b], hs[b]);
  }
}

// one workgroup of ORDER_SCAN threads; thread t owns the ORDER_BINS /
// ORDER_SCAN consecutive positions of the keys in descending order from
// t * ORDER_BINS / ORDER_SCAN, so an exclusive scan over the positions gives
// each key's first position
constexpr int ORDER_SCAN = 1024;
constexpr int ORDER_PER = ORDER_BINS / ORDER_SCAN;
__global__ void __launch_bounds__(ORDER_SCAN) k_order_scan(int* work) {
  __shared__ int2 buf[ORDER_SCAN];
  const int t = threadIdx.x;
  int2 own[ORDER_PER];
  int2 sum = make_int2(0, 0);
#pragma unroll
  for (int i = 0; i < ORDER_PER; i++) {
    const int key = ORDER_BINS - 1 - (t * ORDER_PER + i);
    own[i] = make_int2(work[key], work[ORDER_BINS + key]);
    sum.x += own[i].x;
    sum.y += own[i].y;
  }
  buf[t] = sum;
  __syncthreads();
  for (int d = 1; d < ORDER_SCAN; d <<= 1) {
    int2 v = buf[t];
    if (t >= d) {
      const int2 u = buf[t - d];
      v.x += u.x;
      v.y += u.y;
    }
    __syncthreads();
    buf[t] = v;
    __syncthreads();
  }
  int2 run = make_int2(buf[t].x - sum.x, buf[t].y - sum.y);
#pragma unroll
  for (int i = 0; i < ORDER_PER; i++) {
    const int key = ORDER_BINS - 1 - (t * ORDER_PER + i);
    work[2 * ORDER_BINS + key] = run.x;
    work[3 * ORDER_BINS + key] = run.y;
    run.x += own[i].x;
    run.y += own[i].y;
  }
}

// ORDER_ITEMS columns per thread, one chunk of 256 * ORDER_ITEMS columns per
// workgroup: positions within the chunk from LDS atomics, one global
// reservation per (workgroup, key)
constexpr int ORDER_ITEMS = 4;
__global__ void __launch_bounds__(256) k_order_scatter(int ncol, const int* Lc, const int* Kc,
                                                       int kb, int* order, int* spos,
                                                       int* work) {
  __shared__ int pc[ORDER_BINS], ps[ORDER_BINS];
  for (int b = threadIdx.x; b < ORDER_BINS; b += blockDim.x) pc[b] = ps[b] = 0;
  __syncthreads();
  int key[ORDER_ITEMS], oc[ORDER_ITEMS], os[ORDER_ITEMS], kk[ORDER_ITEMS];
  const int c0 = blockIdx.x * blockDim.x * ORDER_ITEMS + threadIdx.x;
#pragma unroll
  for (int i = 0; i < ORDER_ITEMS; i++) {
    const int c = c0 + i * blockDim.x;
    key[i] = -1;
    if (c < ncol) {
      const int K = Kc[c];
      key[i] = order_key(Lc[c], K, kb);
      kk[i] = max(K, 0);
      oc[i] = atomicAdd(&pc[key[i]], 1);
      os[i] = atomicAdd(&ps[key[i]], kk[i]);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < ORDER_BINS; b += blockDim.x) {
    if (pc[b]) pc[b] = atomicAdd(&work[2 * ORDER_BINS + b], pc[b]);
    if (ps[b]) ps[b] = atomicAdd(&work[3 * ORDER_BINS + b], ps[b]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ORDER_ITEMS; i++) {
    if (key[i] < 0) continue;
    const int c = c0 + i * blockDim.x;
    order[pc[key[i]] + oc[i]] = c;
    if (spos) spos[c] = ps[key[i]] + os[i];  // the column's first solve in the order
  }
}

// solve_order[spos[col] + k] = s for every solve s = toff[col] + k (one thread
// per solve; positions of an inconsistent nk, toff beyond nsolves are skipped)
__global__ void __launch_bounds__(256) k_order_solves(int64_t nsolves, const int* solve_col,
                                                      const int64_t* toff, const int* spos,
                                                      int* solve_order) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsolves) return;
  const int col = solve_col[s];
  const int64_t p = (int64_t)spos[col] + (s - toff[col]);
  if (p >= 0 && p < nsolves) solve_order[p] = (int)s;
}

// ------------------------------------------------------------------ launchers
template <int S>
static hipError_t launch_forward_S(const BatchDev& B, int col_begin, int col_end, int cols_per_wg,
                                   int persistent, int grid, int lds_stride, hipStream_t st) {
  int block = S * cols_per_wg;
  size_t lds = (size_t)cols_per_wg * lds_stride * sizeof(double4);
  if (!persistent) grid = (col_end - col_begin + cols_per_wg - 1) / cols_per_wg;
  if (grid <= 0) return hipSuccess;
  if (lds > 64 * 1024) {
    // dynamic LDS beyond 64 KiB (up to the device's per-workgroup limit)
    hipError_t e = hipFuncSetAttribute((const void*)k_forward<S>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_forward<S>, dim3(grid), dim3(block), lds, st, B, col_begin, col_end,
                     persistent, lds_stride);
  return hipGetLastError();
}

hipError_t launch_forward(const BatchDev& B, int S, int col_begin, int col_end, int cols_per_wg,
                          int persistent, int grid, int lds_stride, hipStream_t st) {
  switch (S) {
    case 8:
      return launch_forward_S<8>(B, col_begin, col_end, cols_per_wg, persistent, grid, lds_stride, st);
    case 16:
      return launch_forward_S<16>(B, col_begin, col_end, cols_per_wg, persistent, grid, lds_stride, st);
    case 32:
      return launch_forward_S<32>(B, col_begin, col_end, cols_per_wg, persistent, grid, lds_stride, st);
    case 64:
      return launch_forward_S<64>(B, col_begin, col_end, cols_per_wg, persistent, grid, lds_stride, st);
    default:
      return hipErrorInvalidValue;
  }
}

int forward_max_blocks(int S, int cols_per_wg, int lds_stride) {
  int block = S * cols_per_wg;
  size_t lds = (size_t)cols_per_wg * lds_stride * sizeof(double4);
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (S) {
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward<8>, block, lds); break;
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward<16>, block, lds); break;
    case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward<32>, block, lds); break;
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward<64>, block, lds); break;
  }
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

hipError_t launch_prep_geom(int nz, const float* depz, float minthk0, int L, void* geo,
                            float* geo_depz, hipStream_t st) {
  hipLaunchKernelGGL(k_prep_geom, dim3((L + 63) / 64), dim3(64), 0, st, nz, depz, minthk0, L,
                     (FlatGeom*)geo, geo_depz);
  return hipGetLastError();
}

hipError_t launch_prep_nodes(int ncol, int nz, const float* vel, const float* vp, const float* rho,
                             const float* depz, const float* geo_depz, int L, const void* geo,
                             double4* layers, double* btp, double* btr, double* cc, float* betmx,
                             int* bad, hipStream_t st) {
  // four columns per 256-thread workgroup (one per wavefront)
  const int grid = (ncol + 3) / 4;
  hipLaunchKernelGGL(k_prep_nodes, dim3(grid), dim3(256), 0, st, ncol, nz, vel, vp, rho, depz,
                     geo_depz, L, (const FlatGeom*)geo, layers, btp, btr, cc, betmx, bad);
  return hipGetLastError();
}

hipError_t launch_prep_layers(int ncol, const int* Lc, const int64_t* loff, const float* thk,
                              const float* vp, const float* vs, const float* rho,
                              const double* thk64, const double* vp64, const double* vs64,
                              const double* rho64, double4* layers, double* btp, double* btr,
                              double* cc, float* betmx, int* bad, hipStream_t st) {
  // four columns per 256-thread workgroup (one per wavefront)
  const int grid = (ncol + 3) / 4;
  hipLaunchKernelGGL(k_prep_layers, dim3(grid), dim3(256), 0, st, ncol, Lc, loff, thk, vp, vs,
                     rho, thk64, vp64, vs64, rho64, layers, btp, btr, cc, betmx, bad);
  return hipGetLastError();
}

hipError_t launch_kernels(const BatchDev& B, const KernDev& O, double* tape, int nq,
                          int grid_threads, hipStream_t st) {
  int block = KERN_BLOCK;
  int grid = std::max(1, grid_threads / block);
  // the checkpoint segments' LDS buffer (beyond the default 64 KiB of
  // dynamic LDS; gfx950 has 160 KiB per workgroup)
  const hipError_t attr = hipFuncSetAttribute(
      (const void*)k_kernels, hipFuncAttributeMaxDynamicSharedMemorySize, KERN_LDS);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(k_kernels, dim3(grid), dim3(block), KERN_LDS, st, B, O, tape, nq);
  return hipGetLastError();
}

hipError_t launch_zero_int(int* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_order(int ncol, int max_layers, const int* Lc, const int* Kc,
                        const int64_t* toff, const int* solve_col, int64_t nsolves, int* order,
                        int* solve_order, int* spos, int* work, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  // period-count bits of the key: as many (up to 4) as leave a key for every
  // layer count up to max_layers
  int kb = 0;
  while (kb < 4 && ((int64_t)(max_layers + 1) << (kb + 1)) <= ORDER_BINS) kb++;
  hipError_t e = launch_zero_int(work, 2 * ORDER_BINS, st);
  if (e != hipSuccess) return e;
  const int hist_grid = std::min((ncol + 255) / 256, 256);
  hipLaunchKernelGGL(k_order_hist, dim3(hist_grid), dim3(256), 0, st, ncol, Lc, Kc, kb, work);
  hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(ORDER_SCAN), 0, st, work);
  const int chunk = 256 * ORDER_ITEMS;
  const bool solves = solve_order && solve_col && nsolves > 0;
  hipLaunchKernelGGL(k_order_scatter, dim3((ncol + chunk - 1) / chunk), dim3(256), 0, st, ncol,
                     Lc, Kc, kb, order, solves ? spos : nullptr, work);
  if (solves)
    hipLaunchKernelGGL(k_order_solves, dim3((unsigned)((nsolves + 255) / 256)), dim3(256), 0, st,
                       nsolves, solve_col, toff, spos, solve_order);
  return hipGetLastError();
}

}  // namespace dk
