// First-arrival traveltimes on the propagation grid: the consumer of the
// phase-velocity maps pvRc in the joint G builder
// (code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:511-725).  A job is one
// (period, source) pair:
//   * the period's (nvz+2) x (nvx+2) control values are diced onto the
//     nnz x nnx propagation grid by cubic B-splines (gridder,
//     CalSurfG.f90:1596-1689): k_gridder, one thread per (period, node);
//   * a refined grid around the source (bsplrefine, CalSurfG.f90:1698-1770)
//     carries the first fast-marching pass until the narrow band touches an
//     inner edge of it (travel with urg = 1, CalSurfG.f90:399-598); its nodes
//     that coincide with coarse nodes seed the band of the coarse pass
//     (CalSurfGAniso_Joint.f90:606-699), which runs to completion (urg = 2);
//   * receiver times come from the coarse field (srtimes,
//     CalSurfG.f90:1772-1906).
//
// Fast marching is a sequential chain (pop the earliest band node, update its
// four neighbours), and its result depends on the order of the pops, so each
// job runs it in the reference's order on one lane.  The parallelism is across
// jobs: one 64-lane workgroup per job slot.  The lanes share the
// data-parallel steps (refined velocities, status clearing, band seeding,
// receivers); in the marching the tree operations run uniformly on the whole
// wave and the up to 16 quadrant solutions of a pop's four neighbours run one
// per lane.  The narrow band (a binary heap with the reference's sift rules)
// lives in LDS for its first HEAP_LDS entries and spills to a per-slot global
// array beyond that.  Each node record holds (traveltime bits, status) in 8
// bytes, so one load fetches both.
//
// Arithmetic is the reference's single precision in its operation order with
// contraction off; division and sqrt are evaluated in double and rounded once
// (the correctly rounded float result).  sin enters only through tables the
// host fills with libm sinf (risti per grid column, sin of each receiver's
// colatitude), as the reference evaluates it.
#include <hip/hip_runtime.h>

#include "dk_internal.h"

namespace dk {
namespace {

constexpr int HEAP_LDS = 1024;
constexpr int FAR = -1;

__device__ __forceinline__ float fdiv(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ float fsqrt(float a) { return (float)sqrt((double)a); }
__device__ __forceinline__ float p2(float x) {
#pragma clang fp contract(off)
  return x * x;
}
__device__ __forceinline__ float p3(float x) {
#pragma clang fp contract(off)
  return x * x * x;
}

// Cubic B-spline basis at parameter (i-1)/n (CalSurfG.f90:1644-1651).
__device__ __forceinline__ void basis(int i, int n, float w[4]) {
#pragma clang fp contract(off)
  float u = (float)n;
  u = fdiv((float)(i - 1), u);
  w[0] = fdiv(p3(1.0f - u), 6.0f);
  w[1] = fdiv(4.0f - 6.0f * p2(u) + 3.0f * p3(u), 6.0f);
  w[2] = fdiv(1.0f + 3.0f * u + 3.0f * p2(u) - 3.0f * p3(u), 6.0f);
  w[3] = fdiv(p3(u), 6.0f);
}

__device__ __forceinline__ float velv(const TravelDev& p, const double* pvk, int i, int j) {
  return (float)pvk[(long)i * (p.nvx + 2) + j];
}

// Cell and in-cell index of a diced node: node s (1-based) of cells of n
// nodes, the last of the ncell cells holding n + 1.
__device__ __forceinline__ void cell_of(int s, int n, int ncell, int* cell, int* in) {
  int c = (s - 1) / n + 1;
  if (c > ncell) c = ncell;
  *cell = c;
  *in = s - n * (c - 1);
}

__global__ void k_gridder(TravelDev p) {
#pragma clang fp contract(off)
  const long nn = (long)p.nnx * p.nnz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn * p.K) return;
  const int k = (int)(t / nn);
  const int node = (int)(t - (long)k * nn);
  const int stz = node % p.nnz + 1, stx = node / p.nnz + 1;
  int i, l, j, m;
  cell_of(stz, p.gdz, p.nvz - 1, &i, &l);
  cell_of(stx, p.gdx, p.nvx - 1, &j, &m);
  float ui[4], vi[4];
  basis(m, p.gdx, ui);
  basis(l, p.gdz, vi);
  const double* pvk = p.pv + (long)k * p.ncol;
  float sumi = 0.0f;
#pragma unroll
  for (int i1 = 1; i1 <= 4; i1++) {
    float sumj = 0.0f;
#pragma unroll
    for (int j1 = 1; j1 <= 4; j1++) sumj = sumj + ui[j1 - 1] * velv(p, pvk, i - 2 + i1, j - 2 + j1);
    sumi = sumi + vi[i1 - 1] * sumj;
  }
  p.veln[t] = sumi;
}

// ---- one grid (coarse or refined) ----
// Node records carry a halo of HALO far nodes on every side, so that the
// records an upwind update reads (up to three nodes from the popped node) are
// always inside the array and load without bounds branches.
constexpr int HALO = 3;

struct Grid {
  int nx, nz, ld;      // nodes; ld = nz + 2*HALO, the records' leading dimension
  float gox, goz, dnx, dnz, earth;
  const float* vel;    // (nz, nx) column-major, no halo
  int2* rec;           // node records (traveltime bits, status), with the halo
  const float* risti;  // earth * sin(colatitude) per column
  int urg, vnl, vnr, vnt, vnb, maxbt, fom;
  __device__ __forceinline__ int lin(int iz, int ix) const {
    return (iz - 1 + HALO) + (ix - 1 + HALO) * ld;
  }
  __device__ __forceinline__ int vlin(int iz, int ix) const { return (iz - 1) + (ix - 1) * nz; }
};

__device__ __forceinline__ int pack(int iz, int ix) { return (ix << 16) | iz; }

// The narrow band: a binary tree of (key bits, packed node) entries with the
// reference's sift rules.  Positions 1..HEAP_LDS are in LDS; deeper ones
// spill to global memory.  The reference also records each node's tree
// position in nsts; here the status record only says far (-1), alive (0) or
// close (1), and updtree finds a close node's position by a search of the
// tree across the wave's lanes, so no tree move writes global memory.
// Every operation comes in two forms: SP = false while the whole tree is in
// LDS (plain LDS accesses), SP = true once it reaches the spill (an access
// that may go to either memory is a flat access, which waits on both the LDS
// and the global memory counters).
struct Heap {
  int2* e;   // LDS: entry p at e[p]
  int2* sp;  // spill: entry p > HEAP_LDS at sp[p - HEAP_LDS - 1]
  int ntr;
  template <bool SP>
  __device__ __forceinline__ int2 get(int p) const {
    const int2 v = (!SP || p <= HEAP_LDS) ? e[p] : sp[p - HEAP_LDS - 1];
    return make_int2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
  }
  template <bool SP>
  __device__ __forceinline__ int node_at(int p) const {
    return (!SP || p <= HEAP_LDS) ? e[p].y : sp[p - HEAP_LDS - 1].y;
  }
  template <bool SP>
  __device__ __forceinline__ void put(int p, int2 v) {
    if (!SP || p <= HEAP_LDS) e[p] = v;
    else sp[p - HEAP_LDS - 1] = v;
  }
};

__device__ __forceinline__ float keyf(int2 v) { return __int_as_float(v.x); }

// addtree / updtree (CalSurfG.f90:879-925, 1005-1032): the entry of node n
// with key tv, starting at position tpc, filters up past larger parents.
template <bool SP>
__device__ __forceinline__ void sift_up_t(Heap& h, int n, float tv, int tpc) {
  int tpp = tpc / 2;
  while (tpp > 0) {
    const int2 par = h.get<SP>(tpp);
    if (!(tv < keyf(par))) break;
    h.put<SP>(tpc, par);
    tpc = tpp;
    tpp = tpc / 2;
  }
  h.put<SP>(tpc, make_int2(__float_as_int(tv), n));
}

__device__ __forceinline__ void sift_up(Heap& h, int n, float tv, int tpc) {
  if (h.ntr <= HEAP_LDS) sift_up_t<false>(h, n, tv, tpc);
  else sift_up_t<true>(h, n, tv, tpc);
}

// downtree (CalSurfG.f90:927-1003): the last entry replaces the root and
// filters down past smaller children.
template <bool SP>
__device__ __forceinline__ void downtree_t(Heap& h) {
  const int2 last = h.get<SP>(h.ntr);
  const float kl = keyf(last);
  h.ntr--;
  const int ntr = h.ntr;
  int tpp = 1, tpc = 2;
  while (tpc < ntr) {
    int2 c = h.get<SP>(tpc);
    const int2 c2 = h.get<SP>(tpc + 1);
    if (keyf(c) > keyf(c2)) {
      tpc++;
      c = c2;
    }
    if (keyf(c) < kl) {
      h.put<SP>(tpp, c);
      tpp = tpc;
      tpc = 2 * tpp;
    } else {
      tpc = ntr + 1;
    }
  }
  if (tpc == ntr) {
    const int2 c = h.get<SP>(tpc);
    if (keyf(c) < kl) {
      h.put<SP>(tpp, c);
      tpp = tpc;
    }
  }
  h.put<SP>(tpp, last);
}

__device__ __forceinline__ void downtree(Heap& h) {
  if (h.ntr == 1) {
    h.ntr = 0;
    return;
  }
  if (h.ntr <= HEAP_LDS) downtree_t<false>(h);
  else downtree_t<true>(h);
}

// Tree position of close node n: every lane compares four entries per round.
template <bool SP>
__device__ __forceinline__ int find_pos_t(const Heap& h, int n) {
  const int lane = threadIdx.x;
  for (int base = 1; base <= h.ntr; base += 256) {
    bool hit[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int p = base + u * 64 + lane;
      hit[u] = p <= h.ntr && h.node_at<SP>(p) == n;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const unsigned long long m = __ballot(hit[u]);
      if (m) return base + u * 64 + __ffsll((long long)m) - 1;
    }
  }
  return 0;
}

// LDS-only form (ntr <= HEAP_LDS): every lane reads its eight entries
// unconditionally (entries past the tree are masked by the position test), so
// the search has no branches until the ballots.  Rounds start at 1, 513, ...,
// so the highest position read is HEAP_LDS, the array's last entry: the eight
// reads share one address with constant offsets.
static_assert(HEAP_LDS % 512 == 0, "find_pos_lds reads positions 1 .. HEAP_LDS");
__device__ __forceinline__ int find_pos_lds(const Heap& h, int n) {
  const int lane = threadIdx.x;
  for (int base = 1; base <= h.ntr; base += 512) {
    int nd[8];
    const int2* row = h.e + base + lane;
#pragma unroll
    for (int u = 0; u < 8; u++) nd[u] = row[u * 64].y;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const unsigned long long m = __ballot(nd[u] == n && base + u * 64 + lane <= h.ntr);
      if (m) return base + u * 64 + __ffsll((long long)m) - 1;
    }
  }
  return 0;
}

__device__ __forceinline__ int find_pos(const Heap& h, int n) {
  return h.ntr <= HEAP_LDS ? find_pos_lds(h, n) : find_pos_t<true>(h, n);
}

__device__ __forceinline__ void set_status(const Grid& g, int packed, int s) {
  g.rec[g.lin(packed & 0xffff, packed >> 16)].y = s;
}

__device__ __forceinline__ float quad_root(float a, float b, float c) {
#pragma clang fp contract(off)
  float rd1 = p2(b) - 4.0f * a * c;
  if (rd1 < 0.0f) rd1 = 0.0f;
  return fdiv(-b + fsqrt(rd1), 2.0f * a);
}

__device__ __forceinline__ float tof(int2 r) { return __int_as_float(r.x); }
__device__ __forceinline__ bool alive(int2 r) { return r.y == 0; }

constexpr float NOSOL = __builtin_inff();

// One quadrant of fouds1 (CalSurfG.f90:608-696): the node's neighbour j
// along its row and k along its column; +inf when neither is alive.
__device__ __forceinline__ float fouds1_quadrant(const Grid& g, float slown, float risti, int2 rj,
                                                 int2 rk) {
#pragma clang fp contract(off)
  const float ri = g.earth;
  float a, b, c, tref;
  if (alive(rj)) {
    if (alive(rk)) {
      const float u = ri * g.dnx, v = risti * g.dnz;
      const float em = tof(rk) - tof(rj);
      a = p2(u) + p2(v);
      b = -2.0f * p2(u) * em;
      c = p2(u) * (p2(em) - p2(v) * p2(slown));
    } else {
      a = 1.0f;
      b = 0.0f;
      c = -p2(slown) * p2(ri) * p2(g.dnx);
    }
    tref = tof(rj);
  } else if (alive(rk)) {
    a = 1.0f;
    b = 0.0f;
    c = -p2(slown * risti * g.dnz);
    tref = tof(rk);
  } else {
    return NOSOL;
  }
  return tref + quad_root(a, b, c);
}

// One quadrant of fouds2 (CalSurfG.f90:698-870): second order along a
// direction whose two upwind nodes (j, j2 or k, k2) are alive and monotone.
__device__ __forceinline__ float fouds2_quadrant(const Grid& g, float slown, float risti, int2 rj,
                                                 int2 rj2, bool vj2, int2 rk, int2 rk2, bool vk2) {
#pragma clang fp contract(off)
  const float ri = g.earth;
  const bool ajz = alive(rj), akx = alive(rk);
  const float tj = tof(rj), tj2 = tof(rj2), tk = tof(rk), tk2 = tof(rk2);
  const bool swj = ajz && vj2 && alive(rj2) && tj > tj2;
  const bool swk = akx && vk2 && alive(rk2) && tk > tk2;
  float a, b, c, tref, tdiv = 1.0f, u, v, em;
  if (swj) {
    if (swk) {
      u = 2.0f * ri * g.dnx;
      v = 2.0f * risti * g.dnz;
      em = 4.0f * tj - tj2 - 4.0f * tk;
      em = em + tk2;
      a = p2(v) + p2(u);
      b = 2.0f * em * p2(u);
      c = p2(u) * (p2(em) - p2(slown) * p2(v));
      tref = 4.0f * tj - tj2;
      tdiv = 3.0f;
    } else if (akx) {
      u = risti * g.dnz;
      v = 2.0f * ri * g.dnx;
      em = 3.0f * tk - 4.0f * tj + tj2;
      a = p2(v) + 9.0f * p2(u);
      b = 6.0f * em * p2(u);
      c = p2(u) * (p2(em) - p2(slown) * p2(v));
      tref = tk;
    } else {
      u = 2.0f * ri * g.dnx;
      a = 1.0f;
      b = 0.0f;
      c = -p2(u) * p2(slown);
      tref = 4.0f * tj - tj2;
      tdiv = 3.0f;
    }
  } else if (ajz) {
    if (swk) {
      u = ri * g.dnx;
      v = 2.0f * risti * g.dnz;
      em = 3.0f * tj - 4.0f * tk + tk2;
      a = p2(v) + 9.0f * p2(u);
      b = 6.0f * em * p2(u);
      c = p2(u) * (p2(em) - p2(v) * p2(slown));
      tref = tj;
    } else if (akx) {
      u = ri * g.dnx;
      v = risti * g.dnz;
      em = tk - tj;
      a = p2(u) + p2(v);
      b = -2.0f * p2(u) * em;
      c = p2(u) * (p2(em) - p2(v) * p2(slown));
      tref = tj;
    } else {
      a = 1.0f;
      b = 0.0f;
      c = -p2(slown) * p2(ri) * p2(g.dnx);
      tref = tj;
    }
  } else if (swk) {
    u = 2.0f * risti * g.dnz;
    a = 1.0f;
    b = 0.0f;
    c = -p2(u) * p2(slown);
    tref = 4.0f * tk - tk2;
    tdiv = 3.0f;
  } else if (akx) {
    a = 1.0f;
    b = 0.0f;
    c = -p2(slown) * p2(risti) * p2(g.dnz);
    tref = tk;
  } else {
    return NOSOL;
  }
  return fdiv(tref + quad_root(a, b, c), tdiv);
}

// bilinear (CalSurfG.f90:2466-2488); nv[i][j] = node (isz-1+j, isx-1+i).
__device__ __forceinline__ float bilinear(const float nv[2][2], float dsx, float dsz, float dnx,
                                          float dnz) {
#pragma clang fp contract(off)
  float biv = 0.0f;
#pragma unroll
  for (int i = 1; i <= 2; i++)
#pragma unroll
    for (int j = 1; j <= 2; j++) {
      const float produ = (1.0f - fabsf(fdiv((float)(i - 1) * dnx - dsx, dnx))) *
                          (1.0f - fabsf(fdiv((float)(j - 1) * dnz - dsz, dnz)));
      biv = biv + nv[i - 1][j - 1] * produ;
    }
  return biv;
}

// Source start (travel, urg /= 2, CalSurfG.f90:423-486): the four nodes of
// the source cell get straight-ray times and enter the band.  Returns 1 for a
// source outside the grid (the reference stops), 2 when the band outgrows the
// tree capacity.  Uniform across the wave.
__device__ __forceinline__ int start_at_source(Heap& h, const Grid& g, float scx, float scz) {
#pragma clang fp contract(off)
  int isx = (int)fdiv(scx - g.gox, g.dnx) + 1;
  int isz = (int)fdiv(scz - g.goz, g.dnz) + 1;
  if (isx < 1 || isx > g.nx || isz < 1 || isz > g.nz) return 1;  // the reference stops
  if (isx == g.nx) isx--;
  if (isz == g.nz) isz--;
  float vss[2][2];
#pragma unroll
  for (int i = 1; i <= 2; i++)
#pragma unroll
    for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = g.vel[g.vlin(isz - 1 + j, isx - 1 + i)];
  const float dsx = (scx - g.gox) - (float)(isx - 1) * g.dnx;
  const float dsz = (scz - g.goz) - (float)(isz - 1) * g.dnz;
  const float vsrc = bilinear(vss, dsx, dsz, g.dnx, g.dnz);
#pragma unroll
  for (int i = 1; i <= 2; i++)
#pragma unroll
    for (int j = 1; j <= 2; j++) {
      const float ds = fsqrt(p2(dsx - (float)(i - 1) * g.dnx) + p2(dsz - (float)(j - 1) * g.dnz));
      const float t = fdiv(2.0f * ds, vss[i - 1][j - 1] + vsrc);
      const int iz = isz - 1 + j, ix = isx - 1 + i;
      g.rec[g.lin(iz, ix)] = make_int2(__float_as_int(t), 1);
      if (h.ntr + 1 > g.maxbt) return 2;
      h.ntr++;
      sift_up(h, pack(iz, ix), t, h.ntr);
    }
  return 0;
}

// The marching loop of travel (CalSurfG.f90:496-597), run by the whole wave:
// the tree operations are uniform (every lane holds the same values and makes
// the same stores, so each lane reads back its own writes), and the upwind
// updates of the up to four neighbours of a popped node are spread over lanes
// 0-15, one quadrant (neighbour n, row side q, column side r) per lane,
// reduced by a minimum over each neighbour's four lanes (the reference's
// running MIN over its quadrants; the times are finite, so the order of the
// minimum does not matter).  Returns 0, or 2 when the band outgrows the tree
// capacity (the reference writes past it).
// FAST: the loop runs while the band has at most HEAP_LDS - 4 entries, so a
// pop (one removal, at most four additions) keeps every tree position in LDS
// and inside the tree capacity (the caller takes this form only when the
// capacity is at least HEAP_LDS): no spill tests and no capacity test.  It
// returns -1 when the band grows past that, and the general form continues.
template <bool REFINED, bool FOM, bool FAST>
__device__ __forceinline__ int march_t(Heap& h, const Grid& g) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  // lane 4n+2q+r: quadrant (row side q, column side r) of neighbour n, the
  // neighbours in the reference's order x-1, x+1, z-1, z+1; record offsets
  // from the popped node
  const int qn = (lane >> 2) & 3, qj = (lane >> 1) & 1, qk = lane & 1;
  const int dxn = qn == 0 ? -1 : qn == 1 ? 1 : 0, dzn = qn == 2 ? -1 : qn == 3 ? 1 : 0;
  const int dj = qj ? 1 : -1, dk = qk ? 1 : -1;
  const int on = dzn + dxn * g.ld;
  const int oj = on + dj * g.ld, oj2 = on + 2 * dj * g.ld, ok = on + dk, ok2 = on + 2 * dk;
  while (h.ntr > 0) {
    if (FAST && h.ntr > HEAP_LDS - 4) return -1;
    const int n0 = h.get<false>(1).y;
    const int ix = n0 >> 16, iz = n0 & 0xffff;
    if (REFINED) {
      const bool edge = (ix == 1 && g.vnl != 1) || (ix == g.nx && g.vnr != g.nx) ||
                        (iz == 1 && g.vnt != 1) || (iz == g.nz && g.vnb != g.nz);
      if (edge) {
        set_status(g, n0, 0);
        return 0;
      }
    }
    const int r0 = g.lin(iz, ix);
    g.rec[r0].y = 0;
    // The updates' records, loaded before the tree work so that their latency
    // overlaps it (downtree touches LDS only); the halo keeps every address in
    // the array, and out-of-grid rows and columns read as far (for j2, k2 the
    // same as the reference's range test).
    const int ixn = ix + dxn, izn = iz + dzn;
    const bool okn = lane < 16 && (unsigned)(ixn - 1) < (unsigned)g.nx &&
                     (unsigned)(izn - 1) < (unsigned)g.nz;
    const bool vj = (unsigned)(ixn + dj - 1) < (unsigned)g.nx;
    const bool vk = (unsigned)(izn + dk - 1) < (unsigned)g.nz;
    const int2 rs = g.rec[r0 + on];
    const int2 rj = g.rec[r0 + oj];
    const int2 rk = g.rec[r0 + ok];
    int2 rj2 = make_int2(0, FAR), rk2 = make_int2(0, FAR);
    if (FOM) {
      rj2 = g.rec[r0 + oj2];
      rk2 = g.rec[r0 + ok2];
    }
    const float vel = g.vel[okn ? g.vlin(izn, ixn) : 0];
    const float risti = g.risti[okn ? ixn - 1 : 0];
    if (FAST) {
      if (h.ntr == 1) h.ntr = 0;
      else downtree_t<false>(h);
    } else {
      downtree(h);
    }
    const int self = okn ? rs.y : 0;  // 0: nothing to update
    float trav = NOSOL;
    if (self != 0 && vj && vk) {
      const float slown = fdiv(1.0f, vel);
      trav = FOM ? fouds2_quadrant(g, slown, risti, rj, rj2, true, rk, rk2, true)
                 : fouds1_quadrant(g, slown, risti, rj, rk);
    }
    // minimum over each neighbour's quad of lanes (DPP quad permutations
    // [1,0,3,2] and [2,3,0,1])
    trav = fminf(trav, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(trav), 0xB1, 0xF,
                                                               0xF, false)));
    trav = fminf(trav, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(trav), 0x4E, 0xF,
                                                               0xF, false)));
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int s = __builtin_amdgcn_readlane(self, 4 * q);  // 0 (skip) off the grid
      if (s == 0) continue;
      const float t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(trav), 4 * q));
      const int izq = iz + (q == 2 ? -1 : q == 3 ? 1 : 0);
      const int ixq = ix + (q == 0 ? -1 : q == 1 ? 1 : 0);
      const int nq = pack(izq, ixq);
      g.rec[g.lin(izq, ixq)] = make_int2(__float_as_int(t), 1);
      if (FAST) {
        if (s == FAR) {
          h.ntr++;
          sift_up_t<false>(h, nq, t, h.ntr);
        } else {
          sift_up_t<false>(h, nq, t, find_pos_lds(h, nq));
        }
      } else if (s == FAR) {
        if (h.ntr + 1 > g.maxbt) return 2;
        h.ntr++;
        sift_up(h, nq, t, h.ntr);
      } else {
        sift_up(h, nq, t, find_pos(h, nq));
      }
    }
  }
  return 0;
}

template <bool REFINED, bool FOM>
__device__ __forceinline__ int march(Heap& h, const Grid& g) {
  if (g.maxbt >= HEAP_LDS) {
    const int rc = march_t<REFINED, FOM, true>(h, g);
    if (rc >= 0) return rc;
  }
  return march_t<REFINED, FOM, false>(h, g);
}

__device__ __forceinline__ int nint_f(float x) { return (int)roundf(x); }

// srtimes (CalSurfG.f90:1772-1906) for one receiver on the coarse grid.
__device__ __forceinline__ int receiver_time(const TravelDev& p, const Grid& c, float scx, float scz, float rcx,
                             float rcz, float srx, float* out) {
#pragma clang fp contract(off)
  int irx = (int)fdiv(rcx - c.gox, c.dnx) + 1;
  int irz = (int)fdiv(rcz - c.goz, c.dnz) + 1;
  if (irx == c.nx) irx--;
  if (irz == c.nz) irz--;
  const int isx = (int)fdiv(scx - c.gox, c.dnx) + 1;
  const int isz = (int)fdiv(scz - c.goz, c.dnz) + 1;
  float sred = p2((scx - rcx) * c.earth);
  sred = sred + p2((scz - rcz) * c.earth * srx);
  sred = fsqrt(sred);
  const bool sw = sred < p.dpl || (isx == irx && isz == irz);
  float trr;
  if (sw) {
    if (isx >= c.nx || isz >= c.nz) return 4;
    float vss[2][2];
#pragma unroll
    for (int k = 1; k <= 2; k++)
#pragma unroll
      for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = c.vel[c.vlin(isz - 1 + l, isx - 1 + k)];
    float drx = (scx - c.gox) - (float)(isx - 1) * c.dnx;
    float drz = (scz - c.goz) - (float)(isz - 1) * c.dnz;
    const float vels = bilinear(vss, drx, drz, c.dnx, c.dnz);
#pragma unroll
    for (int k = 1; k <= 2; k++)
#pragma unroll
      for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = c.vel[c.vlin(irz - 1 + l, irx - 1 + k)];
    drx = (rcx - c.gox) - (float)(irx - 1) * c.dnx;
    drz = (rcz - c.goz) - (float)(irz - 1) * c.dnz;
    const float velr = bilinear(vss, drx, drz, c.dnx, c.dnz);
    trr = fdiv(2.0f * sred, vels + velr);
  } else {
    const float drx = (rcx - c.gox) - (float)(irx - 1) * c.dnx;
    const float drz = (rcz - c.goz) - (float)(irz - 1) * c.dnz;
    trr = 0.0f;
#pragma unroll
    for (int k = 1; k <= 2; k++)
#pragma unroll
      for (int l = 1; l <= 2; l++) {
        const float produ = (1.0f - fabsf(fdiv((float)(l - 1) * c.dnz - drz, c.dnz))) *
                            (1.0f - fabsf(fdiv((float)(k - 1) * c.dnx - drx, c.dnx)));
        trr = trr + tof(c.rec[c.lin(irz - 1 + l, irx - 1 + k)]) * produ;
      }
  }
  *out = trr;
  return 0;
}

template <bool FOM>
__global__ __launch_bounds__(64) void k_travel(TravelDev p) {
#pragma clang fp contract(off)
  __shared__ int2 hent[HEAP_LDS + 1];
  __shared__ int s_rc;
  const int lane = threadIdx.x;
  const long slot = blockIdx.x;
  const long nn = (long)p.nnx * p.nnz;
  const long ncp = (long)(p.nnx + 2 * HALO) * (p.nnz + 2 * HALO);
  int2* recc = p.node_c + slot * ncp;
  int2* recr = p.node_r + slot * p.nrp_max;
  float* velr = p.vel_r + slot * (long)p.nr_max;
  Heap h = {hent, p.spill + slot * p.spill_cap, 0};

  Grid c;
  c.nx = p.nnx;
  c.nz = p.nnz;
  c.ld = p.nnz + 2 * HALO;
  c.gox = p.gox;
  c.goz = p.goz;
  c.dnx = p.dnx;
  c.dnz = p.dnz;
  c.earth = p.earth;
  c.rec = recc;
  c.risti = p.risti_c;
  c.urg = 2;
  c.vnl = c.vnr = c.vnt = c.vnb = 0;
  c.maxbt = p.maxbt_c;
  c.fom = p.fom;

  for (long q = slot; q < p.njob; q += gridDim.x) {
    const int kk = p.job_k[q];
    const float scx = p.scx[q], scz = p.scz[q];
    const double* pvk = p.pv + (long)kk * p.ncol;
    c.vel = p.veln + (long)kk * nn;
    int rc = 0;
    if (p.asgr) {
      int isx = (int)fdiv(scx - p.gox, p.dnx) + 1, isz = (int)fdiv(scz - p.goz, p.dnz) + 1;
      if (isx == p.nnx) isx--;
      if (isz == p.nnz) isz--;
      const int vnl = max(isx - p.sgs, 1), vnr = min(isx + p.sgs, p.nnx);
      const int vnt = max(isz - p.sgs, 1), vnb = min(isz + p.sgs, p.nnz);
      Grid r;
      r.nx = (vnr - vnl) * p.sgdl + 1;
      r.nz = (vnb - vnt) * p.sgdl + 1;
      r.ld = r.nz + 2 * HALO;
      r.dnx = p.drnx;
      r.dnz = p.drnz;
      r.gox = p.gox + p.dnx * (float)(vnl - 1);
      r.goz = p.goz + p.dnz * (float)(vnt - 1);
      r.earth = p.earth;
      r.vel = velr;
      r.rec = recr;
      r.risti = p.risti_r + (long)(vnl - 1) * p.nrx_max;
      r.urg = 1;
      r.vnl = vnl;
      r.vnr = vnr;
      r.vnt = vnt;
      r.vnb = vnb;
      r.fom = p.fom;
      r.maxbt = (r.nx > p.nnx || r.nz > p.nnz)
                    ? nint_f(p.snb * (float)max(r.nx, p.nnx) * (float)max(r.nz, p.nnz))
                    : p.maxbt_c;
      // refined velocities (bsplrefine) and far status, all lanes
      const int nrzr = p.gdz * p.sgdl, nrxr = p.gdx * p.sgdl;
      const int origz = (vnt - 1) * p.sgdl + 1, origx = (vnl - 1) * p.sgdl + 1;
      const int nr = r.nx * r.nz;
      for (int e = lane; e < nr; e += 64) {
        const int idm1 = e % r.nz + 1, idm2 = e / r.nz + 1;
        int i, kq, j, l;
        cell_of(origz + idm1 - 1, nrzr, p.nvz - 1, &i, &kq);
        cell_of(origx + idm2 - 1, nrxr, p.nvx - 1, &j, &l);
        float ui[4], vi[4], sum[4];
        basis(l, nrxr, ui);
        basis(kq, nrzr, vi);
#pragma unroll
        for (int i1 = 1; i1 <= 4; i1++) {
          sum[i1 - 1] = 0.0f;
#pragma unroll
          for (int j1 = 1; j1 <= 4; j1++)
            sum[i1 - 1] = sum[i1 - 1] + ui[j1 - 1] * velv(p, pvk, i - 2 + i1, j - 2 + j1);
          sum[i1 - 1] = vi[i1 - 1] * sum[i1 - 1];
        }
        velr[e] = sum[0] + sum[1] + sum[2] + sum[3];
      }
      for (int e = lane; e < (r.nx + 2 * HALO) * r.ld; e += 64) recr[e] = make_int2(0, FAR);
      __syncthreads();
      h.ntr = 0;
      rc = start_at_source(h, r, scx, scz);
      if (rc == 0) rc = march<true, FOM>(h, r);
      if (lane == 0) s_rc = rc;
      for (long e = lane; e < ncp; e += 64) recc[e] = make_int2(0, FAR);
      __syncthreads();
      rc = s_rc;
      // coincident nodes seed the coarse pass (CalSurfGAniso_Joint.f90:606-630)
      const int bx = vnr - vnl + 1, bz = vnb - vnt + 1;
      if (rc == 0)
        for (int e = lane; e < bx * bz; e += 64) {
          const int a = e % bz, b = e / bz;
          const int2 v = recr[r.lin(a * p.sgdl + 1, b * p.sgdl + 1)];
          int2& w = recc[c.lin(vnt + a, vnl + b)];
          if (v.y >= 0) w = v;
          else w.y = v.y;
        }
      __syncthreads();
      // alive nodes next to far nodes rejoin the band (:663-683)
      if (rc == 0)
        for (int e = lane; e < bx * bz; e += 64) {
          const int l = vnt + e % bz, k = vnl + e / bz;
          if (recc[c.lin(l, k)].y != 0) continue;
          const bool far = (l - 1 >= 1 && recc[c.lin(l - 1, k)].y == FAR) ||
                           (l + 1 <= c.nz && recc[c.lin(l + 1, k)].y == FAR) ||
                           (k - 1 >= 1 && recc[c.lin(l, k - 1)].y == FAR) ||
                           (k + 1 <= c.nx && recc[c.lin(l, k + 1)].y == FAR);
          if (far) recc[c.lin(l, k)].y = 1;
        }
      __syncthreads();
      if (rc == 0) {
        // travel with urg = 2: the band in the reference's scan order (x
        // outer), the whole wave uniformly
        h.ntr = 0;
        c.maxbt = r.maxbt;
        for (int b = 0; b < bx && rc == 0; b++)
          for (int a = 0; a < bz; a++) {
            const int iz = vnt + a, ix = vnl + b;
            const int2 v = recc[c.lin(iz, ix)];
            if (v.y > 0) {
              if (h.ntr + 1 > c.maxbt) {
                rc = 2;
                break;
              }
              h.ntr++;
              sift_up(h, pack(iz, ix), tof(v), h.ntr);
            }
          }
        if (rc == 0) rc = march<false, FOM>(h, c);
        if (lane == 0) s_rc = rc;
      }
    } else {
      for (long e = lane; e < ncp; e += 64) recc[e] = make_int2(0, FAR);
      __syncthreads();
      h.ntr = 0;
      c.urg = 0;
      c.maxbt = p.maxbt_c;
      rc = start_at_source(h, c, scx, scz);
      if (rc == 0) rc = march<false, FOM>(h, c);
      if (lane == 0) s_rc = rc;
    }
    __syncthreads();
    rc = s_rc;
    if (rc == 0 && p.asgr) {
      // The refined pass stops at the refined grid's far x (z) edge unless
      // vnr (vnb), a coarse index, equals the refined node count
      // (CalSurfG.f90:503-523).  For a source close to the model's south or
      // east edge that happens before the band reaches a coarse node; the
      // coarse band then starts empty and the reference leaves ttn unset.
      bool unreached = false;
      for (long e = lane; e < nn; e += 64)
        unreached |= recc[c.lin((int)(e % p.nnz) + 1, (int)(e / p.nnz) + 1)].y != 0;
      if (unreached) atomicMax(&s_rc, 3);
      __syncthreads();
      rc = s_rc;
    }
    const long r0 = p.rec_off[q], r1 = p.rec_off[q + 1];
    int bad = 0;
    for (long e = r0 + lane; e < r1; e += 64) {
      float t = __int_as_float(0x7fc00000);
      if (rc == 0 && receiver_time(p, c, scx, scz, p.rcx[e], p.rcz[e], p.sin_rcx[e], &t)) bad = 4;
      p.times[e] = t;
    }
    if (p.ttn)
      for (long e = lane; e < nn; e += 64)
        p.ttn[q * nn + e] = tof(recc[c.lin((int)(e % p.nnz) + 1, (int)(e / p.nnz) + 1)]);
    // a receiver read past the grid (source on the last node) marks the job
    if (bad) atomicMax(&s_rc, 4);
    __syncthreads();
    if (lane == 0) p.status[q] = s_rc;
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_travel(const TravelDev& p, hipStream_t st) {
  const long n = (long)p.nnx * p.nnz * p.K;
  if (n > 0)
    hipLaunchKernelGGL(k_gridder, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p);
  if (p.njob > 0 && p.slots > 0) {
    if (p.fom)
      hipLaunchKernelGGL(k_travel<true>, dim3((unsigned)p.slots), dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL(k_travel<false>, dim3((unsigned)p.slots), dim3(64), 0, st, p);
  }
  return hipGetLastError();
}

}  // namespace dk
