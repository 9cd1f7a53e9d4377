// G-row assembly of the joint isotropic + azimuthal inversion: the sparse
// rows (rw, iw, col) that CalSurfGAniso_Joint builds from each ray's path
// weights and the stage outputs (CalSurfGAniso_Joint.f90:750-797).
//
// For a ray (data row iw) of period k with path weights fdm, fdmc, fdms on the
// interior cells (jj, kk), jj = 1..nvz, kk = 1..nvx, and column
// post = jj*(nvx+2) + kk of the stage's nx*ny grid, the reference fills a
// dense row of 3 * nparpi entries, nparpi = nvx*nvz*(nz-1):
//   Vs block: real(eff(post, k, z) * fdm)     eff = (sen_vp coe_a + sen_rho coe_rho) + sen_vs
//   Gc block: Lsen_Gsc(post, k, z) * fdmc     (float32)
//   Gs block: Lsen_Gsc(post, k, z) * fdms     (float32)
// at index nn = block*nparpi + z*nvx*nvz + (jj-1)*nvx + kk for every cell with
// |fdm| >= ftol, and keeps the entries with |row(nn)| > ftol in increasing nn.
//
// Here a ray arrives as its list of touched cells (ascending cell index
// (jj-1)*nvx + kk-1, with their fdm, fdmc, fdms).  A first kernel forms the
// effective dc/dVs (the same float64 expression as k_effective) and copies
// Lsen_Gsc into depth-fastest planes effT, lsenT (period, column, depth), so
// that one cell's depth run is contiguous.  One wave per ray then works with
// lanes on cells: each lane holds its cell's weights and plane offset and
// reads its depth run a 16-byte vector at a time (the loads, not the bytes,
// bound this gather, so fewer and wider loads are what counts).  The count
// pass sums the kept entries per lane; a tiled scan turns the row counts into
// offsets; the fill pass walks the row in the reference's order (block,
// depth, cell), compacts each step's kept entries with a ballot and writes rw
// and col, then fills the row's iw range.  Rays of more than 128 cells, or
// depth counts that are not a multiple of 4, take the same order with scalar
// loads.  Workgroups are mapped to XCDs so that the rays in flight on one XCD
// share one or two periods (their planes stay in that XCD's L2).
#include <hip/hip_runtime.h>

#include "dk_brocher.h"
#include "dk_internal.h"

namespace dk {

struct GrowArgs {
  int nvx, nvz, nz, K, ncol;
  long nray;
  int row_base;
  float ftol;
  const int* ray_k;
  const long* cell_off;
  const int* cell;
  const float* fdm;
  const float* fdmc;
  const float* fdms;
  const double* effT;
  const float* lsenT;
  long cap;
  float* rw;
  int* iw;
  int* col;
  long* row_nnz;
  const long* row_off;
};

constexpr int GROW_WAVES = 4;   // rays per workgroup, one wave each
constexpr int NXCD = 8;         // MI355X: workgroups are dealt to the 8 XCDs round robin

// effT[(k*ncol + col)*nd + z] = eff(z, k, col), lsenT likewise from Lsen_Gsc.
// k_grow_planes: one workgroup per (64-column tile, period); the tile is read
// along columns (coalesced), transposed in LDS (row stride nd + 1: the
// column-strided LDS writes fall in distinct banks) and written as one
// contiguous run.  k_grow_planes_wide: one thread per (column, period), for
// depth counts whose tile does not fit in LDS.
constexpr int PT = 64;  // columns per planes tile

__global__ void __launch_bounds__(256) k_grow_planes(int ncol, int nz, int K,
                                                     const float* __restrict__ vel,
                                                     const double* __restrict__ svs,
                                                     const double* __restrict__ svp,
                                                     const double* __restrict__ srho,
                                                     const float* __restrict__ lsen,
                                                     double* __restrict__ effT,
                                                     float* __restrict__ lsenT) {
  extern __shared__ double pe[];  // PT x (nd + 1) doubles, then PT x (nd + 1) floats
  const int nd = nz - 1, rs = nd + 1;
  float* pl = (float*)(pe + PT * rs);
  const int col0 = blockIdx.x * PT, k = blockIdx.y;
  const int nc = min(PT, ncol - col0);
  for (int idx = threadIdx.x; idx < nd * PT; idx += blockDim.x) {
    const int z = idx / PT, c = idx - z * PT;
    if (c < nc) {
      const int col = col0 + c;
      float ca, cr;
      brocher_coeffs(vel[(long)z * ncol + col], &ca, &cr);
      const long o = ((long)z * K + k) * ncol + col;
      pe[c * rs + z] = effective_vs(svs[o], svp[o], srho[o], ca, cr);
      pl[c * rs + z] = lsen[o];
    }
  }
  __syncthreads();
  const long base = ((long)k * ncol + col0) * nd;
  for (int idx = threadIdx.x; idx < nc * nd; idx += blockDim.x) {
    const int c = idx / nd, z = idx - c * nd;
    effT[base + idx] = pe[c * rs + z];
    lsenT[base + idx] = pl[c * rs + z];
  }
}

__global__ void k_grow_planes_wide(int ncol, int nz, int K, const float* __restrict__ vel,
                                   const double* __restrict__ svs, const double* __restrict__ svp,
                                   const double* __restrict__ srho,
                                   const float* __restrict__ lsen, double* __restrict__ effT,
                                   float* __restrict__ lsenT) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ncol * K) return;
  const int col = (int)(t % ncol), k = (int)(t / ncol);
  const int nd = nz - 1;
  for (int z = 0; z < nd; z++) {
    float ca, cr;
    brocher_coeffs(vel[(long)z * ncol + col], &ca, &cr);
    const long o = ((long)z * K + k) * ncol + col;
    effT[t * nd + z] = effective_vs(svs[o], svp[o], srho[o], ca, cr);
    lsenT[t * nd + z] = lsen[o];
  }
}

struct RayCtx {
  long c0;
  int k, nc, nd, cells, nparpi;
};

// This lane's cell for chunk ch of the ray: validity, cell index, weights and
// the offset of its depth run in the planes.
struct CellReg {
  bool ok;
  int ci;
  float w, wc, ws;
  long base;
};

__device__ inline CellReg grow_cell(const GrowArgs& a, const RayCtx& R, int i) {
  CellReg c;
  const bool in = i < R.nc;
  const long o = R.c0 + (in ? i : 0);
  c.ci = in ? a.cell[o] : 0;
  c.w = in ? a.fdm[o] : 0.0f;
  c.wc = in ? a.fdmc[o] : 0.0f;
  c.ws = in ? a.fdms[o] : 0.0f;
  // an out-of-range cell gives no entry (the host entry rejects it)
  c.ok = in && c.ci >= 0 && c.ci < R.cells && fabsf(c.w) >= a.ftol;
  const int cc = min(max(c.ci, 0), R.cells - 1);
  const int jj = cc / a.nvx;
  const int post = (jj + 1) * (a.nvx + 2) + (cc - jj * a.nvx) + 1;
  c.base = ((long)(in ? R.k : 0) * a.ncol + post) * R.nd;
  return c;
}

// Depth tile widths of the vector loads: 2 doubles (16 B) of effT, 4 floats of
// lsenT; usable when the depth runs keep that alignment (nd a multiple of 4).
constexpr int TE = 2, TL = 4;

template <bool WRITE>
__global__ void __launch_bounds__(64 * GROW_WAVES) k_grows(GrowArgs a, int nblk) {
  const int lane = threadIdx.x & 63;
  // XCD-aware order: XCD x (blocks x, x+8, ...) takes the x-th contiguous share
  // of the rays, so the rays in flight on one XCD share one or two periods and
  // their planes stay in that XCD's L2
  const int per = nblk / NXCD;
  const int lb = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  const long r = (long)lb * GROW_WAVES + (threadIdx.x >> 6);
  if (r >= a.nray) return;  // whole waves only: no workgroup barrier is used
  RayCtx R;
  R.c0 = a.cell_off[r];
  R.k = a.ray_k[r];
  // a ray with an invalid period gets an empty row (the host entry rejects it)
  R.nc = (R.k >= 0 && R.k < a.K) ? (int)(a.cell_off[r + 1] - R.c0) : 0;
  R.nd = a.nz - 1;
  R.cells = a.nvx * a.nvz;
  R.nparpi = R.cells * R.nd;
  const int nd = R.nd, nc = R.nc;
  const int nch = (nc + 63) >> 6;
  const bool vec = (nd % TL) == 0;
  long pos = WRITE ? a.row_off[r] : 0;
  if (!WRITE) {
    // count: the order does not matter; each lane runs down its cell's depths
    for (int ch = 0; ch < nch; ch++) {
      const CellReg c = grow_cell(a, R, ch * 64 + lane);
      int n = 0;
      if (c.ok) {
        if (vec) {
          const double2* e2 = (const double2*)(a.effT + c.base);
          const float4* l4 = (const float4*)(a.lsenT + c.base);
          for (int z = 0; z < nd / TE; z++) {
            const double2 e = e2[z];
            n += fabsf((float)(e.x * (double)c.w)) > a.ftol;
            n += fabsf((float)(e.y * (double)c.w)) > a.ftol;
          }
          for (int z = 0; z < nd / TL; z++) {
            const float4 l = l4[z];
            n += (fabsf(l.x * c.wc) > a.ftol) + (fabsf(l.y * c.wc) > a.ftol) +
                 (fabsf(l.z * c.wc) > a.ftol) + (fabsf(l.w * c.wc) > a.ftol);
            n += (fabsf(l.x * c.ws) > a.ftol) + (fabsf(l.y * c.ws) > a.ftol) +
                 (fabsf(l.z * c.ws) > a.ftol) + (fabsf(l.w * c.ws) > a.ftol);
          }
        } else {
          for (int z = 0; z < nd; z++) {
            n += fabsf((float)(a.effT[c.base + z] * (double)c.w)) > a.ftol;
            const float l = a.lsenT[c.base + z];
            n += (fabsf(l * c.wc) > a.ftol) + (fabsf(l * c.ws) > a.ftol);
          }
        }
      }
      // wave sum of the lanes' counts
      for (int d = 32; d > 0; d >>= 1) n += __shfl_xor(n, d);
      pos += n;
    }
    if (lane == 0) a.row_nnz[r] = pos;
    return;
  }
  const long pos0 = pos;
  // rw and col per kept entry; iw, constant along the row, is filled below
  auto emit = [&](bool keep, float v, int nn) {
    const unsigned long long m = __ballot(keep);
    if (keep) {
      const long p = pos + __popcll(m & ((1ull << lane) - 1ull));
      if (p < a.cap) {
        a.rw[p] = v;
        a.col[p] = nn;
      }
    }
    pos += __popcll(m);
  };
  if (nch <= 2 && vec) {
    // fill in the reference's order (block, depth, cell); lanes hold cells of
    // up to two chunks and read their depth runs a vector at a time
    CellReg c[2];
    c[0] = grow_cell(a, R, lane);
    c[1] = grow_cell(a, R, 64 + lane);
    for (int z0 = 0; z0 < nd; z0 += TE) {
      double2 e[2];
#pragma unroll
      for (int ch = 0; ch < 2; ch++)
        if (ch < nch) e[ch] = *(const double2*)(a.effT + c[ch].base + z0);
#pragma unroll
      for (int j = 0; j < TE; j++)
#pragma unroll
        for (int ch = 0; ch < 2; ch++)
          if (ch < nch) {
            const float v = (float)((j == 0 ? e[ch].x : e[ch].y) * (double)c[ch].w);
            emit(c[ch].ok && fabsf(v) > a.ftol, v, (z0 + j) * R.cells + c[ch].ci + 1);
          }
    }
#pragma unroll
    for (int b = 1; b < 3; b++)
      for (int z0 = 0; z0 < nd; z0 += TL) {
        float4 l[2];
#pragma unroll
        for (int ch = 0; ch < 2; ch++)
          if (ch < nch) l[ch] = *(const float4*)(a.lsenT + c[ch].base + z0);
#pragma unroll
        for (int j = 0; j < TL; j++)
#pragma unroll
          for (int ch = 0; ch < 2; ch++)
            if (ch < nch) {
              const float lv = j == 0 ? l[ch].x : j == 1 ? l[ch].y : j == 2 ? l[ch].z : l[ch].w;
              const float v = lv * (b == 1 ? c[ch].wc : c[ch].ws);
              emit(c[ch].ok && fabsf(v) > a.ftol, v,
                   b * R.nparpi + (z0 + j) * R.cells + c[ch].ci + 1);
            }
      }
  } else {
    for (int b = 0; b < 3; b++)
      for (int z = 0; z < nd; z++)
        for (int ch = 0; ch < nch; ch++) {
          const CellReg c = grow_cell(a, R, ch * 64 + lane);
          float v;
          if (b == 0)
            v = (float)(a.effT[c.base + z] * (double)c.w);
          else
            v = a.lsenT[c.base + z] * (b == 1 ? c.wc : c.ws);
          emit(c.ok && fabsf(v) > a.ftol, v, b * R.nparpi + z * R.cells + c.ci + 1);
        }
  }
  const long hi = pos < a.cap ? pos : a.cap;
  for (long p = pos0 + lane; p < hi; p += 64) a.iw[p] = a.row_base + (int)r;
}

// Exclusive scan of the per-ray counts in tiles of SCAN_TILE: k_scan_tiles
// scans each tile and writes its sum; k_scan_add adds the sum of the tiles
// before it (each block sums those few tile totals itself) and the last block
// writes the total.
constexpr int SCAN_TILE = 4096;  // 1024 threads x 4 counts

__global__ void __launch_bounds__(1024) k_scan_tiles(long n, const long* __restrict__ cnt,
                                                     long* __restrict__ off,
                                                     long* __restrict__ tsum) {
  __shared__ long part[1024];
  const int t = threadIdx.x;
  const long lo = (long)blockIdx.x * SCAN_TILE + 4L * t;
  long v[4];
  long s = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[q] = lo + q < n ? cnt[lo + q] : 0;
    s += v[q];
  }
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const long x = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  long run = part[t] - s;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (lo + q < n) off[lo + q] = run;
    run += v[q];
  }
  if (t == 1023) tsum[blockIdx.x] = part[1023];
}

__global__ void __launch_bounds__(256) k_scan_add(long n, int ntiles, long* __restrict__ off,
                                                  const long* __restrict__ tsum,
                                                  long* __restrict__ total) {
  __shared__ long red[256];
  const int t = threadIdx.x;
  const int tile = blockIdx.x;
  long s = 0;
  for (int i = t; i < tile; i += 256) s += tsum[i];
  red[t] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (t < d) red[t] += red[t + d];
    __syncthreads();
  }
  const long pre = red[0];
  const long lo = (long)tile * SCAN_TILE;
  for (int i = t; i < SCAN_TILE; i += 256)
    if (lo + i < n) off[lo + i] += pre;
  if (tile == ntiles - 1 && t == 0) total[0] = pre + tsum[tile];
}

// workgroups for nray rays: a multiple of the XCD count
static unsigned grow_blocks(long nray) {
  const long nb = (nray + GROW_WAVES - 1) / GROW_WAVES;
  return (unsigned)((nb + NXCD - 1) / NXCD * NXCD);
}

long grow_scan_tiles(long nray) { return (nray + SCAN_TILE - 1) / SCAN_TILE; }

hipError_t launch_grows(const GrowDev& g, hipStream_t st) {
  GrowArgs a;
  a.nvx = g.nvx;
  a.nvz = g.nvz;
  a.nz = g.nz;
  a.K = g.K;
  a.ncol = (g.nvx + 2) * (g.nvz + 2);
  a.nray = g.nray;
  a.row_base = g.row_base;
  a.ftol = g.ftol;
  a.ray_k = g.ray_k;
  a.cell_off = (const long*)g.cell_off;
  a.cell = g.cell;
  a.fdm = g.fdm;
  a.fdmc = g.fdmc;
  a.fdms = g.fdms;
  a.effT = g.effT;
  a.lsenT = g.lsenT;
  a.cap = g.cap;
  a.rw = g.rw;
  a.iw = g.iw;
  a.col = g.col;
  a.row_nnz = (long*)g.row_nnz;
  a.row_off = (const long*)g.row_off;
  if (g.nray == 0) {
    // the total of an empty batch: one tile of no counts
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, 0L, (const long*)g.row_nnz,
                       (long*)g.row_off, (long*)g.tsum);
    hipLaunchKernelGGL(k_scan_add, dim3(1), dim3(256), 0, st, 0L, 1, (long*)g.row_off,
                       (const long*)g.tsum, (long*)g.nar);
    return hipGetLastError();
  }
  const size_t plds = (size_t)PT * g.nz * 12;  // PT x (nd + 1) x (8 + 4) bytes
  if (plds <= 48 * 1024) {
    hipLaunchKernelGGL(k_grow_planes, dim3((unsigned)((a.ncol + PT - 1) / PT), (unsigned)g.K),
                       dim3(256), plds, st, a.ncol, g.nz, g.K, g.vel, g.svs, g.svp, g.srho,
                       g.lsen, g.effT, g.lsenT);
  } else {
    const long npl = (long)a.ncol * g.K;
    hipLaunchKernelGGL(k_grow_planes_wide, dim3((unsigned)((npl + 255) / 256)), dim3(256), 0, st,
                       a.ncol, g.nz, g.K, g.vel, g.svs, g.svp, g.srho, g.lsen, g.effT, g.lsenT);
  }
  const unsigned nb = grow_blocks(g.nray);
  hipLaunchKernelGGL(k_grows<false>, dim3(nb), dim3(64 * GROW_WAVES), 0, st, a, (int)nb);
  const int nt = (int)grow_scan_tiles(g.nray);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(1024), 0, st, (long)g.nray,
                     (const long*)g.row_nnz, (long*)g.row_off, (long*)g.tsum);
  hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(256), 0, st, (long)g.nray, nt, (long*)g.row_off,
                     (const long*)g.tsum, (long*)g.nar);
  hipLaunchKernelGGL(k_grows<true>, dim3(nb), dim3(64 * GROW_WAVES), 0, st, a, (int)nb);
  return hipGetLastError();
}

}  // namespace dk
