// HIP kernels of the TI (dc/dA, dc/dL -> Lsen_Gsc) part of the stage.
//
//   k_prep_ti   node model -> refined layers (refineLayerMdl), TI moduli
//               TA = TC = rho Vp^2, TL = TN = rho Vs^2, TF = TA - 2 TL, their
//               earth flattening and sphericity factors; one thread per column
//               (depthkernelTI.f90:96-138, tregn96.f:774-841, :1301-1383).
//   k_ti        one thread per (column, period): tregn96's eigenfunctions and
//               kernels at the float32 phase velocity of the root search, and
//               the float32 splice into Lsen_Gsc(post, period, interval)
//               (depthkernelTI.f90:139-162), in real arithmetic.  A solve with
//               a layer whose P/SV eigenvalues form a complex-conjugate pair is
//               appended to a queue instead.
//   k_ti_cx     the queued solves in complex arithmetic (tregn96's complex*16
//               form), in a kernel of their own so that k_ti keeps its
//               register budget.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dk_geom.h"
#include "dk_internal.h"
#include "dk_ti.h"

namespace dk {

static_assert(TI_SCRATCH == ti::Q_N, "TI scratch doubles per layer");
static_assert(TI_SCRATCH_CX == ti::Q_NC, "complex TI scratch doubles per layer");

__device__ float brocher_vp_ti(float v) {
#pragma clang fp contract(off)
  float v2 = v * v, v3 = v2 * v, v4 = v3 * v;
  return 0.9409f + 2.0947f * v - 0.8206f * v2 + 0.2683f * v3 - 0.0251f * v4;
}
__device__ float brocher_rho_ti(float p) {
#pragma clang fp contract(off)
  float p2 = p * p, p3 = p2 * p, p4 = p3 * p, p5 = p4 * p;
  return 1.6612f * p - 0.4721f * p2 + 0.0671f * p3 - 0.0043f * p4 + 0.000106f * p5;
}

// The flattened layer of a record: TC = TA, TN = TL (depthkernelTI.f90:124-138),
// without its constants.
__device__ __forceinline__ ti::Layer layer_of(const TiLayerDev& t) {
  ti::Layer r;
  r.d = (double)t.df;
  r.A = (double)t.A;
  r.C = (double)t.A;
  r.F = (double)t.F;
  r.L = (double)t.L;
  r.N = (double)t.L;
  r.rho = (double)t.rho;
  return r;
}

// One refined layer of the TI model: moduli (depthkernelTI.f90:124-138), TI
// flattening with radius 6371 (sphere_tdisp96, tregn96.f:774-841) and the
// sphericity factor with radius 6370 (bldsph, tregn96.f:1301-1383).  The
// flattening and sphericity factors depend on the layer thicknesses only, so
// they are computed once per node geometry (ti_geom in dk_geom.h, run by
// k_prep_geom) and every column applies them (ti_layer).
__device__ __forceinline__ TiLayerDev ti_layer(const TiGeom& g, float rvp, float rvs, float rrho) {
#pragma clang fp contract(off)
  TiLayerDev t;
  t.d = g.d;
  t.rvp = rvp;
  t.rvs = rvs;
  t.rrho = rrho;
  float v2 = t.rvp * t.rvp, s2 = t.rvs * t.rvs;
  t.TA = t.rrho * v2;
  t.TL = t.rrho * s2;
  float tl2 = 2.0f * t.TL;
  t.TF = 1.0f * (t.TA - tl2);
  t.rho = (float)((double)t.rrho * g.pr);
  t.A = (float)((double)t.TA * g.fm);
  t.L = (float)((double)t.TL * g.fm);
  t.F = t.TF;
  t.df = g.df;
  t.vtp = g.vtp;
  t.pad = 0.0f;
  ti::Layer r = layer_of(t);
  ti::layer_constants(&r);
  t.k[0] = r.iC;
  t.k[1] = r.iL;
  t.k[2] = r.eta;
  t.k[3] = r.FiC;
  t.k[4] = r.ah;
  t.k[5] = r.av;
  t.k[6] = r.bv;
  t.k[7] = r.Feta;
  return t;
}

// The records are stored in word planes: 8-byte word w of the record of
// (layer m, column col) at [(m * TI_WORDS + w) * ncol + col], so that the
// lanes of a wavefront (consecutive columns) read a word of their records in
// one contiguous access.  Words: 0 (d, rvp), 1 (rvs, rrho), 2 (TA, TL),
// 3 (TF, df), 4 (A, L), 5 (F, rho), 6 (vtp, pad), 7..14 k[0..7].
constexpr int TI_WORDS = (int)(sizeof(TiLayerDev) / 8);
static_assert(sizeof(TiLayerDev) == 8 * TI_WORDS, "TI record words");
union TiWords {
  TiLayerDev t;
  uint64_t u[TI_WORDS];
};
__device__ __forceinline__ uint64_t ti_word(const TiLayerDev* recs, int ncol, int col, int m,
                                            int w) {
  return reinterpret_cast<const uint64_t*>(recs)[((size_t)m * TI_WORDS + w) * ncol + col];
}

// Node models, one thread per (layer, column), columns fastest: each thread
// interpolates its layer between its nodes (the sublayer mapping and TI
// factors of geo) and writes its record's words.
__global__ void __launch_bounds__(256) k_prep_ti(int ncol, const float* __restrict__ vel,
                                                 const float* __restrict__ vpin,
                                                 const float* __restrict__ rhoin, int L,
                                                 const FlatGeom* __restrict__ geo,
                                                 TiLayerDev* __restrict__ out,
                                                 int* __restrict__ queue) {
#pragma clang fp contract(off)
  auto node_vals = [&](int col, int i, float* vs, float* vp, float* rho) {
    const size_t o = (size_t)i * ncol + col;
    *vs = vel[o];
    if (vpin) {
      *vp = vpin[o];
      *rho = rhoin[o];
    } else {
      *vp = brocher_vp_ti(*vs);
      *rho = brocher_rho_ti(*vp);
    }
  };
  uint64_t* words = reinterpret_cast<uint64_t*>(out);
  const size_t n = (size_t)L * ncol;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    if (t == 0) queue[0] = 0;  // the complex-pair queue of the k_ti launch that follows
    const int m = (int)(t / ncol), col = (int)(t % ncol);
    const FlatGeom& g = geo[m];
    if (g.node < 0) continue;
    float vs, vp, rho;
    node_vals(col, g.node, &vs, &vp, &rho);
    if (m < L - 1) {
      float vs_n, vp_n, rho_n;
      node_vals(col, g.node + 1, &vs_n, &vp_n, &rho_n);
      const float den = (float)(2 * g.ns);
      const float w = (float)(2 * g.sub - 1);
      vp = vp + w * (vp_n - vp) / den;
      vs = vs + w * (vs_n - vs) / den;
      rho = rho + w * (rho_n - rho) / den;
    }
    TiWords r;
    r.t = ti_layer(g.ti, vp, vs, rho);
#pragma unroll
    for (int wd = 0; wd < TI_WORDS; wd++) words[((size_t)m * TI_WORDS + wd) * ncol + col] = r.u[wd];
  }
}

// One (column, period) solve s of the TI splice: writes Lsen_Gsc of every
// interval (when lsen is set) and, when dcdA/dcdL are set, the chained
// kernels dc/dA, dc/dL of every refined layer m at ((m*K + k)*ncol + col);
// zeros where no kernel exists (no phase velocity, vanishing surface minor),
// or NaN for a missing phase velocity with nan_failed.  Returns solve<G_>'s
// code; on -2 nothing is written.
template <class G_, bool KL>
__device__ int ti_one(long s, int ncol, int K, int L, int nz, const TiLayerDev* __restrict__ layers,
                      const double* __restrict__ periods, const double* __restrict__ pv,
                      const int* __restrict__ nsub, float* __restrict__ lsen, float* __restrict__ dcdA,
                      float* __restrict__ dcdL, double* scratch, long stride, int nan_failed) {
  const float twopi = 2.0f * 3.141592654f;
  const int col = (int)(s % ncol), k = (int)(s / ncol);
  auto W = [&](int m, int w) { return ti_word(layers, ncol, col, m, w); };
  auto out = [&](int j) -> float& { return lsen[(size_t)col + (size_t)ncol * ((size_t)k + (size_t)K * j)]; };
  auto lay_out = [&](float* a, int m) -> float& {
    return a[(size_t)col + (size_t)ncol * ((size_t)k + (size_t)K * m)];
  };
  auto fill_all = [&](float v) {
    if (lsen)
      for (int j = 0; j < nz - 1; j++) out(j) = v;
    if constexpr (KL)
      for (int m = 0; m < L; m++) {
        lay_out(dcdA, m) = v;
        lay_out(dcdL, m) = v;
      }
  };
  const double c = (double)(float)pv[s];
  if (!(c > 0.0)) {
    // no phase velocity for this period (root search failed): no kernel
    fill_all(nan_failed ? __builtin_nanf("") : 0.0f);
    return 0;
  }
  const double omega = (double)twopi / (double)(float)periods[k];
  auto head = [&](int m) {
    TiWords q;
    q.u[3] = W(m, 3);
    q.u[4] = W(m, 4);
    q.u[5] = W(m, 5);
    q.u[7] = W(m, 7);
    q.u[8] = W(m, 8);
    ti::Layer r = layer_of(q.t);
    r.iC = q.t.k[0];
    r.iL = q.t.k[1];
    return r;
  };
  auto tail = [&](int m, ti::Layer* r) {
    TiWords q;
#pragma unroll
    for (int w = 9; w < TI_WORDS; w++) q.u[w] = W(m, w);
    r->eta = q.t.k[2];
    r->FiC = q.t.k[3];
    r->ah = q.t.k[4];
    r->av = q.t.k[5];
    r->bv = q.t.k[6];
    r->Feta = q.t.k[7];
  };
  double isc, inrm;
  const int rc = ti::solve<G_>(L, head, tail, omega, c, scratch, stride, &isc, &inrm);
  if (rc == -1) fill_all(0.0f);
  if (rc != 0) return rc;
  // tregn96's kernels top layer first, chained to dc/dA, dc/dL and spliced into
  // the intervals; the factors and record of layer m+1 are loaded while layer
  // m is chained (one memory latency per layer instead of one per load)
  using T = typename G_::T;
  struct Rec {
    double f0, f1, f2;
    float vtp, rrho, rvp, rvs, TA, TL, TF;
  };
  auto load = [&](int m) {
    Rec r;
    r.f0 = ti::factor<T>(scratch, stride, m, 0);
    r.f1 = ti::factor<T>(scratch, stride, m, 1);
    r.f2 = ti::factor<T>(scratch, stride, m, 2);
    TiWords q;
    q.u[0] = W(m, 0);
    q.u[1] = W(m, 1);
    q.u[2] = W(m, 2);
    q.u[3] = W(m, 3);
    q.u[6] = W(m, 6);
    r.vtp = q.t.vtp;
    r.rrho = q.t.rrho;
    r.rvp = q.t.rvp;
    r.rvs = q.t.rvs;
    r.TA = q.t.TA;
    r.TL = q.t.TL;
    r.TF = q.t.TF;
    return r;
  };
  // the half-space feeds no interval; without KL the loop stops at the last
  // interval
  const int mend = KL ? L : (nz > 1 ? L - 1 : 0);
  int interval = 0, sub = 0;
  int ns = nz > 1 ? nsub[0] : 0;
  float acc = 0.0f;
  Rec nx = load(0);
  for (int m = 0; m < mend; m++) {
    const Rec r = nx;
    nx = load(m + 1 < L ? m + 1 : L - 1);
    const double vt = (double)r.vtp;
    const float dah = ti::kernel_f32(r.f0 * vt * isc), dbv = ti::kernel_f32(r.f1 * vt * isc),
                dn = ti::kernel_f32(r.f2 * inrm);
    float dA, dL;
    ti::chain_AL(r.rrho, r.rvp, r.rvs, r.TA, r.TL, r.TF, dah, dbv, dn, &dA, &dL);
    if constexpr (KL) {
      lay_out(dcdA, m) = dA;
      lay_out(dcdL, m) = dL;
      if (!lsen) continue;
    }
    if (m >= L - 1 || interval >= nz - 1) continue;
    acc = ti::splice_add(acc, r.TA, r.TL, dA, dL);
    if (++sub == ns) {
      out(interval) = acc;
      acc = 0.0f;
      sub = 0;
      interval++;
      ns = nsub[interval < nz - 1 ? interval : nz - 2];
    }
  }
  return 0;
}

// KL: also write the per-layer kernels dcdA, dcdL (a variant of its own, so
// that the splice-only kernel keeps its register budget of two waves per SIMD)
template <bool KL>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
k_ti(int ncol, int K, int L, int nz,
                                            const TiLayerDev* __restrict__ layers,
                                            const double* __restrict__ periods,
                                            const double* __restrict__ pv,
                                            const int* __restrict__ nsub, float* __restrict__ lsen,
                                            float* __restrict__ dcdA, float* __restrict__ dcdL,
                                            double* __restrict__ scratch, int* __restrict__ queue,
                                            int nan_failed) {
  const long nthr = (long)gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const long nsolves = (long)ncol * K;
  for (long s = tid; s < nsolves; s += nthr) {
    const int rc = ti_one<ti::Eig, KL>(s, ncol, K, L, nz, layers, periods, pv, nsub, lsen, dcdA,
                                       dcdL, scratch + tid, nthr, nan_failed);
    if (rc == -2) queue[1 + atomicAdd(queue, 1)] = (int)s;
  }
}

template <bool KL>
__global__ void __launch_bounds__(64) k_ti_cx(int ncol, int K, int L, int nz,
                                              const TiLayerDev* __restrict__ layers,
                                              const double* __restrict__ periods,
                                              const double* __restrict__ pv,
                                              const int* __restrict__ nsub,
                                              float* __restrict__ lsen,
                                              float* __restrict__ dcdA,
                                              float* __restrict__ dcdL,
                                              double* __restrict__ scratch,
                                              const int* __restrict__ queue) {
  const long nthr = (long)gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = queue[0];
  for (long i = tid; i < n; i += nthr)
    ti_one<ti::EigC, KL>(queue[1 + i], ncol, K, L, nz, layers, periods, pv, nsub, lsen, dcdA,
                         dcdL, scratch + tid, nthr, 0);
}

hipError_t launch_prep_ti(int ncol, const float* vel, const float* vp, const float* rho, int L,
                          const void* geo, TiLayerDev* out, int* queue, hipStream_t st) {
  // one thread per (layer, column)
  const size_t n = (size_t)L * ncol;
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_prep_ti, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, ncol, vel, vp, rho,
                     L, (const FlatGeom*)geo, out, queue);
  return hipGetLastError();
}

int ti_cx_threads(int nthr) {
  // the complex solves reuse k_ti's scratch: TI_SCRATCH_CX * n <= TI_SCRATCH * nthr
  int n = (int)((long)nthr * TI_SCRATCH / TI_SCRATCH_CX) / 64 * 64;
  return n < 64 ? 64 : n;
}

hipError_t launch_ti(int ncol, int K, int L, int nz, const TiLayerDev* layers,
                     const double* periods, const double* pv, const int* nsub, float* lsen,
                     float* dcdA, float* dcdL, double* scratch, int nthr, int* queue,
                     int nan_failed, hipStream_t st) {
  int block = 128;
  int grid = nthr / block;
  if (grid < 1) grid = 1;
  if (dcdA)
    hipLaunchKernelGGL(k_ti<true>, dim3(grid), dim3(block), 0, st, ncol, K, L, nz, layers, periods,
                       pv, nsub, lsen, dcdA, dcdL, scratch, queue, nan_failed);
  else
    hipLaunchKernelGGL(k_ti<false>, dim3(grid), dim3(block), 0, st, ncol, K, L, nz, layers,
                       periods, pv, nsub, lsen, dcdA, dcdL, scratch, queue, nan_failed);
  if (hipError_t e = hipGetLastError()) return e;
  // the queued complex-pair solves; the grid is sized for the worst case and
  // exits at once when the queue is empty
  int ncx = ti_cx_threads(nthr);
  int cap = ncol * K;
  if (ncx > (cap + 63) / 64 * 64) ncx = (cap + 63) / 64 * 64;
  if (dcdA)
    hipLaunchKernelGGL(k_ti_cx<true>, dim3(ncx / 64), dim3(64), 0, st, ncol, K, L, nz, layers,
                       periods, pv, nsub, lsen, dcdA, dcdL, scratch, (const int*)queue);
  else
    hipLaunchKernelGGL(k_ti_cx<false>, dim3(ncx / 64), dim3(64), 0, st, ncol, K, L, nz, layers,
                       periods, pv, nsub, lsen, dcdA, dcdL, scratch, (const int*)queue);
  return hipGetLastError();
}

// grad_A(m, col) = sum_k grad_c(k, col) dc_k/dA_m, and the same for L, over the
// refined layers of node models: one thread per (layer, column), periods in
// order (deterministic).  dcdA, dcdL (L, K, ncol) f32; grad_c (K, ncol) f64;
// grad_A, grad_L (L, ncol) f64.  Periods with a non-finite kernel (NaN at a
// missing phase velocity) or a zero grad_c contribute nothing.
__global__ void k_ti_vjp(int ncol, int L, int K, const float* __restrict__ dcdA,
                         const float* __restrict__ dcdL, const double* __restrict__ gc,
                         double* __restrict__ gA, double* __restrict__ gL) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ncol * L) return;
  const int m = (int)(t / ncol), col = (int)(t % ncol);
  double a = 0.0, l = 0.0;
  for (int k = 0; k < K; k++) {
    const double g = gc[(long)k * ncol + col];
    const long o = ((long)m * K + k) * ncol + col;
    const double da = (double)dcdA[o], dl = (double)dcdL[o];
    if (g != 0.0 && isfinite(da) && isfinite(dl)) {
      a = fma(g, da, a);
      l = fma(g, dl, l);
    }
  }
  if (gA) gA[t] = a;
  if (gL) gL[t] = l;
}

hipError_t launch_ti_vjp(int ncol, int L, int K, const float* dcdA, const float* dcdL,
                         const double* gc, double* gA, double* gL, hipStream_t st) {
  const long n = (long)ncol * L;
  hipLaunchKernelGGL(k_ti_vjp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ncol, L, K,
                     dcdA, dcdL, gc, gA, gL);
  return hipGetLastError();
}

}  // namespace dk
