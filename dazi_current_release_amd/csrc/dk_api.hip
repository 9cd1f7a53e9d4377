// Host side of the C ABI declared in include/dazi_dk.h.
//
// Device memory for intermediate data (flattened models, roots, the adjoint
// tape, the work counter) is owned by a per-device pool that grows on demand
// and is reused across calls; the caller owns inputs and outputs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <memory>
#include <vector>

#include "../../include/dazi_dk.h"
#include "dk_internal.h"

namespace {

thread_local std::string g_err;
// cols_per_wg 0: chosen per launch (resolve_shape)
dazi_dk_tune g_tune = {0, 0, 0, 64, 2.0e-6};
int g_timing = 0;

int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err(DAZI_DK_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

enum Slot {
  S_LAYERS,
  S_BTP,  // btp, then btr = btp^-2.275, one double each per layer
  S_CC,
  S_BETMX,
  S_BAD,
  S_CROOT,
  S_STATUS,
  S_WORK,
  S_TAPE,
  S_ORDER,   // ragged batches: columns in scheduling order, their first solve positions,
             // then the sort's counters
  S_SORDER,  // ragged batches: solves in scheduling order
  S_H_VEL,
  S_H_VP,
  S_H_RHO,
  S_H_DEPZ,
  S_H_T,
  S_H_PV,
  S_H_SVS,
  S_H_SVP,
  S_H_SRHO,
  S_TI_LAYERS,
  S_TI_SCRATCH,
  S_TI_QUEUE,
  S_TISTAT,
  S_H_LSEN,
  S_G_RAYK,
  S_G_CELLOFF,
  S_G_CELL,
  S_G_W,
  S_G_CNT,
  S_G_ROWOFF,
  S_G_NAR,
  S_G_RW,
  S_G_IW,
  S_G_COL,
  S_G_EFF,
  S_G_LSENT,
  S_G_TSUM,
  S_A_G,
  S_A_T,
  S_A_OUT,
  S_COUNT
};

// One node geometry: the depth values and minthk it was derived from, its
// refined layer count L, the sublayer count of every node interval (device,
// nz - 1 ints) and the layer records that every column shares (device: L
// FlatGeom records, then the nz depths they were built from).
struct Geometry {
  std::vector<float> depz;
  float minthk = 0.0f;
  int L = 0;
  void* nsub = nullptr;
  void* recs = nullptr;
  bool built = false;  // k_prep_geom has been enqueued for recs
};

struct Pool {
  int dev = -1;
  void* p[S_COUNT] = {};
  size_t n[S_COUNT] = {};
  int ncu = 0;
  hipEvent_t ev[8] = {};
  // side stream for work that runs beside the caller's stream (the output DMA
  // of the host entry points), joined back with events; default priority
  hipStream_t side = nullptr;
  hipEvent_t fk[4] = {};
  int nforward = 0;
  int last_L = 0, last_ti_threads = 0;  // refined layers and TI threads of the last call
  bool have_events = false;
  // every node geometry met on this device (never freed: a captured graph
  // may read any of them), and the one of the last call that took depz as a
  // device pointer, keyed by that pointer, nz and minthk
  std::vector<std::unique_ptr<Geometry>> geos;
  Geometry* cur = nullptr;
  const float* cur_depz = nullptr;
  int cur_nz = 0;
  float cur_minthk = 0.0f;
  size_t geo_bytes = 0;
  std::vector<void*> retired;  // buffers replaced by growth (see pool_get)
  size_t retired_bytes = 0;
  // bumped by every call that may overwrite the stage's work slots; a host
  // call may reuse the root search of the previous host call only while it is
  // unchanged
  uint64_t gen = 0;
  // the previous host entry point call: its inputs, launch configuration and
  // the pool generation it left (HostKey)
  struct HostKey {
    bool valid = false;
    uint64_t gen = 0;
    int nx = 0, ny = 0, nz = 0, kmax = 0, mode = 0;
    float minthk = 0.0f;
    dazi_dk_tune tune = {};
    std::vector<float> vel, vp, rho, depz;
    std::vector<double> t;
  } hkey;
  // pinned staging of the host entry points' outputs (two buffers: the DMA
  // of one chunk overlaps the host copy of the previous)
  void* stage_buf[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {};
  bool stage_init = false;
};

Pool g_pool[16];

int get_pool(Pool** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err(DAZI_DK_ENODEV, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= 16) return set_err(DAZI_DK_ENODEV, "device index %d unsupported", dev);
  Pool* P = &g_pool[dev];
  if (P->dev < 0) {
    P->dev = dev;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    P->ncu = prop.multiProcessorCount;
    for (int i = 0; i < 8; i++) HIPCHK(hipEventCreate(&P->ev[i]));
    HIPCHK(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
    for (int i = 0; i < 4; i++) HIPCHK(hipEventCreateWithFlags(&P->fk[i], hipEventDisableTiming));
  }
  *out = P;
  return 0;
}

// A slot only grows (by at least half its size).  The buffer it replaces may
// still be read by work in flight or by a captured HIP graph, so it is retired
// rather than freed: retired buffers live until the process ends; since each
// growth is at least 1.5x, a slot's retired buffers add up to at most twice
// its final size.
int pool_get(Pool* P, Slot s, size_t bytes, void** out) {
  if (bytes == 0) bytes = 8;
  if (P->n[s] < bytes) {
    size_t nb = std::max(bytes, P->n[s] + P->n[s] / 2);
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, nb));
    if (P->p[s]) {
      P->retired.push_back(P->p[s]);
      P->retired_bytes += P->n[s];
    }
    P->p[s] = q;
    P->n[s] = nb;
  }
  *out = P->p[s];
  return 0;
}

// Refined-layer geometry of the node model (CalSurfG.f90:2506-2520), float32.
int node_geometry(int nz, const float* depz, float minthk0, std::vector<int>* nsub) {
  nsub->resize(nz > 1 ? nz - 1 : 0);
  int L = 0;
  for (int i = 0; i < nz - 1; i++) {
    volatile float thk = depz[i + 1] - depz[i];
    volatile float minthk = thk / minthk0;
    volatile float q = (thk + 1.0e-4f) / minthk;
    int ns = (int)q + 1;
    if (!(thk > 0.0f) || ns < 1) return -1;
    (*nsub)[i] = ns;
    L += ns;
  }
  return L + 1;
}

// lanes_per_scan x cols_per_wg, once both are known: whole waves, at most 256
int check_block(int S, int cpw) {
  const int block = S * cpw;
  if (S != 0 && cpw != 0 && (block > 256 || (block % 64) != 0))
    return set_err(DAZI_DK_EINVAL,
                   "lanes_per_scan*cols_per_wg must be a multiple of 64 and <= 256 (got %d*%d)",
                   S, cpw);
  return 0;
}

int check_tune(const dazi_dk_tune* t, dazi_dk_tune* out) {
  *out = t ? *t : g_tune;
  if (out->lanes_per_scan == 0) out->lanes_per_scan = g_tune.lanes_per_scan;
  if (out->cols_per_wg == 0) out->cols_per_wg = g_tune.cols_per_wg;
  if (out->tile_columns == 0) out->tile_columns = 64;
  if (!(out->refine_tol > 0.0)) out->refine_tol = 2.0e-6;
  if (out->refine_tol < 1.0e-12) out->refine_tol = 1.0e-12;
  int S = out->lanes_per_scan;
  if (S != 0 && S != 8 && S != 16 && S != 32 && S != 64)
    return set_err(DAZI_DK_EINVAL, "lanes_per_scan must be 0, 8, 16, 32 or 64 (got %d)", S);
  if (out->cols_per_wg < 0)
    return set_err(DAZI_DK_EINVAL, "cols_per_wg must be >= 0 (got %d)", out->cols_per_wg);
  if (int rc = check_block(S, out->cols_per_wg)) return rc;
  if (out->mode != 0 && out->mode != 1) return set_err(DAZI_DK_EINVAL, "mode must be 0 or 1");
  if (out->tile_columns < 1) return set_err(DAZI_DK_EINVAL, "tile_columns must be >= 1");
  return 0;
}

void rec(Pool* P, int i, hipStream_t st) {
  if (g_timing) (void)hipEventRecord(P->ev[i], st);
}

// LDS per workgroup of the current device (160 KiB on gfx950), looked up once
// per device.
size_t lds_per_block() {
  static int lim[16] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 64 * 1024;
  if (lim[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 64 * 1024;
    lim[dev] = v;
  }
  return (size_t)lim[dev];
}

// The root search's workgroup shape for ncol columns of up to lmax layers on
// ncu compute units.  lanes_per_scan = 0 takes the most lanes per column, of
// 16, 32 and 64, that keep the launch within one wave per SIMD (16 when even
// those exceed it): a small batch then scans more trial velocities per round
// and needs fewer rounds, a large one keeps the wave count at the device's
// SIMD count or above.  The LDS staging holds two double4 records per layer
// and column.  cols_per_wg = 0 picks 256-thread workgroups (one wave on each
// SIMD of a CU: smaller groups can leave the waves of a one-wave-per-SIMD
// launch crowded on part of the device), halved while the layer records do
// not fit.
int resolve_shape(dazi_dk_tune* T, int lmax, int64_t ncol, int ncu) {
  if (T->lanes_per_scan == 0) {
    const int64_t simd = 4 * (int64_t)(ncu > 0 ? ncu : 1);
    int S = 64;
    while (S > 16 && ncol * S / 64 > simd) S /= 2;
    T->lanes_per_scan = S;
    if (int rc = check_block(S, T->cols_per_wg)) return rc;
  }
  const size_t per_col = (size_t)2 * lmax * 32, lim = lds_per_block();
  if (T->cols_per_wg == 0) {
    int cpw = 256 / T->lanes_per_scan;
    while (T->lanes_per_scan * cpw > 64 && (size_t)cpw * per_col > lim) cpw /= 2;
    T->cols_per_wg = cpw;
  }
  if ((size_t)T->cols_per_wg * per_col > lim)
    return set_err(DAZI_DK_ELIMIT,
                   "model too deep for LDS staging: %d layers x %d columns need %zu bytes of LDS "
                   "(the device has %zu per workgroup)",
                   lmax, T->cols_per_wg, (size_t)T->cols_per_wg * per_col, lim);
  return 0;
}

// Root search over columns [0, ncol) with the configured mode on a device of
// ncu compute units; *nlaunch receives the number of launches.  The persistent
// work counter B.work is cleared by a kernel (no memset node in a graph).
int forward_launches(const dk::BatchDev& B, const dazi_dk_tune& T0, int lmax, int ncu,
                     hipStream_t st, int* nlaunch) {
  dazi_dk_tune T = T0;
  if (int rc = resolve_shape(&T, lmax, B.ncol, ncu)) return rc;
  int S = T.lanes_per_scan, cpw = T.cols_per_wg;
  int lds_stride = 2 * lmax;
  *nlaunch = 0;
  if (T.mode == 0) {
    HIPCHK(dk::launch_zero_int(B.work, 1, st));
    int nb = dk::forward_max_blocks(S, cpw, lds_stride);
    int grid = ncu * nb;
    int need = (B.ncol + cpw - 1) / cpw;
    if (grid > need) grid = need;
    HIPCHK(dk::launch_forward(B, S, 0, B.ncol, cpw, 1, grid, lds_stride, st));
    *nlaunch = 1;
  } else {
    // (64-bit bounds: a tile of the whole batch may be given as a huge count)
    for (int64_t c0 = 0; c0 < B.ncol; c0 += T.tile_columns) {
      const int64_t c1 = std::min<int64_t>(c0 + T.tile_columns, B.ncol);
      HIPCHK(dk::launch_forward(B, S, (int)c0, (int)c1, cpw, 0, 0, lds_stride, st));
      ++*nlaunch;
    }
  }
  return 0;
}

int run_forward(Pool* P, const dk::BatchDev& B, const dazi_dk_tune& T, int lmax,
                hipStream_t st) {
  P->nforward = 0;
  return forward_launches(B, T, lmax, P->ncu, st, &P->nforward);
}

int kernel_threads(Pool* P, int64_t nsolves) {
  int64_t cap = (int64_t)P->ncu * 512;
  int64_t n = nsolves < cap ? nsolves : cap;
  n = (n + 255) / 256 * 256;
  if (n < 256) n = 256;
  return (int)n;
}

int ti_threads(Pool* P, int64_t nsolves) {
  // two waves per SIMD (k_ti fits two in the register file)
  int64_t cap = (int64_t)P->ncu * 512;
  int64_t n = nsolves < cap ? nsolves : cap;
  n = (n + 127) / 128 * 128;
  if (n < 128) n = 128;
  return (int)n;
}

// Refined geometry (L and the per-interval sublayer counts on the device).  With
// depz_host NULL, depz_dev is read back once per (pointer, nz, minthk) and the
// result is reused by later calls, which then enqueue kernels only (no host
// synchronisation: such calls can be captured in a HIP graph).
// The node geometry of a call (L, sublayer counts, shared layer records),
// keyed on the values of depz and minthk, not on the address of depz.  A call
// on device pointers whose (depz pointer, nz, minthk) matches the previous
// such call reads depz back and compares its values, unless the caller
// declares them unchanged (DAZI_DK_DEPZ_UNCHANGED) or the stream is being
// captured (no host synchronisation is possible there); the prep kernels then
// compare depz with the geometry's own copy on the device and mark every
// column of a changed depz with status DK_STATUS_GEOMETRY.  A geometry met
// before is reused with its buffers; a new one gets buffers of its own.
int geometry(Pool* P, const dazi_dk_desc* d, const float* depz_dev, const float* depz_host,
             hipStream_t st, Geometry** out) {
  std::vector<float> dz;
  const bool from_dev = depz_host == nullptr;
  if (from_dev && P->cur && P->cur_depz == depz_dev && P->cur_nz == d->nz &&
      P->cur_minthk == d->minthk0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    bool same = (d->flags & DAZI_DK_DEPZ_UNCHANGED) || cs != hipStreamCaptureStatusNone;
    if (!same) {
      dz.resize(d->nz);
      HIPCHK(hipMemcpyAsync(dz.data(), depz_dev, sizeof(float) * d->nz, hipMemcpyDeviceToHost,
                            st));
      HIPCHK(hipStreamSynchronize(st));
      same = std::memcmp(dz.data(), P->cur->depz.data(), sizeof(float) * d->nz) == 0;
    }
    if (same) {
      *out = P->cur;
      return 0;
    }
  }
  if (from_dev) {
    if (dz.empty()) {
      dz.resize(d->nz);
      HIPCHK(hipMemcpyAsync(dz.data(), depz_dev, sizeof(float) * d->nz, hipMemcpyDeviceToHost,
                            st));
      HIPCHK(hipStreamSynchronize(st));
    }
    depz_host = dz.data();
  }
  Geometry* G = nullptr;
  for (auto& g : P->geos)
    if (g->minthk == d->minthk0 && g->depz.size() == (size_t)d->nz &&
        std::memcmp(g->depz.data(), depz_host, sizeof(float) * d->nz) == 0) {
      G = g.get();
      break;
    }
  if (!G) {
    std::vector<int> nsub;
    int Lc = node_geometry(d->nz, depz_host, d->minthk0, &nsub);
    if (Lc < 2) return set_err(DAZI_DK_EINVAL, "depz must be strictly increasing");
    if (Lc > 4096) return set_err(DAZI_DK_ELIMIT, "refined model has %d layers (limit 4096)", Lc);
    auto g = std::make_unique<Geometry>();
    g->depz.assign(depz_host, depz_host + d->nz);
    g->minthk = d->minthk0;
    g->L = Lc;
    const size_t nb = nsub.size() * 4, rb = (size_t)Lc * dk::PREP_GEOM_BYTES + (size_t)d->nz * 4;
    HIPCHK(hipMalloc(&g->nsub, nb));
    if (hipMalloc(&g->recs, rb) != hipSuccess) {
      (void)hipFree(g->nsub);
      return set_err(DAZI_DK_EHIP, "hipMalloc of %zu bytes of layer records failed", rb);
    }
    P->geo_bytes += nb + rb;
    Geometry* gp = g.get();
    P->geos.push_back(std::move(g));
    HIPCHK(hipMemcpyAsync(gp->nsub, nsub.data(), nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // nsub is a host temporary
    G = gp;
  }
  if (from_dev) {
    P->cur = G;
    P->cur_depz = depz_dev;
    P->cur_nz = d->nz;
    P->cur_minthk = d->minthk0;
  } else {
    P->cur = nullptr;
  }
  *out = G;
  return 0;
}

// A small pool of host threads that splits large host memcpys (the staged
// copy-out of the host entry points), created on first use.
class CopyThreads {
 public:
  static constexpr int kWorkers = 3;
  ~CopyThreads() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void copy(void* dst, const void* src, size_t n) {
    if (n < ((size_t)1 << 20)) {
      std::memcpy(dst, src, n);
      return;
    }
    start();
    const size_t part = (n / (kWorkers + 1) + 63) / 64 * 64;
    {
      std::lock_guard<std::mutex> g(m_);
      dst_ = (char*)dst;
      src_ = (const char*)src;
      n_ = n;
      part_ = part;
      pending_ = kWorkers;
      epoch_++;
    }
    cv_.notify_all();
    std::memcpy(dst, src, std::min(part, n));
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  void start() {
    if (!th_.empty()) return;
    for (int w = 0; w < kWorkers; w++) th_.emplace_back([this, w] { run(w + 1); });
  }
  void run(int slice) {
    uint64_t seen = 0;
    while (true) {
      char* d;
      const char* s;
      size_t n, part;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || epoch_ != seen; });
        if (stop_) return;
        seen = epoch_;
        d = dst_;
        s = src_;
        n = n_;
        part = part_;
      }
      const size_t off = (size_t)slice * part;
      if (off < n) std::memcpy(d + off, s + off, std::min(part, n - off));
      {
        std::lock_guard<std::mutex> g(m_);
        if (--pending_ == 0) done_cv_.notify_all();
      }
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, part_ = 0;
  int pending_ = 0;
  uint64_t epoch_ = 0;
  bool stop_ = false;
};
CopyThreads g_copy;

constexpr size_t kStageChunk = (size_t)8 << 20;

// Copies `bytes` of device memory at src to the caller's host memory at dst
// after everything enqueued on `st`: chunks go by DMA (the side stream) into
// the two pinned staging buffers and from there by host threads to dst, the
// DMA of chunk c+1 overlapping the host copy of chunk c.  Synchronous.
int copy_out(Pool* P, void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return 0;
  if (!P->stage_init) {
    for (int i = 0; i < 2; i++) {
      HIPCHK(hipHostMalloc(&P->stage_buf[i], kStageChunk, hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&P->stage_ev[i], hipEventDisableTiming));
    }
    P->stage_init = true;
  }
  HIPCHK(hipEventRecord(P->fk[3], st));
  HIPCHK(hipStreamWaitEvent(P->side, P->fk[3], 0));
  const size_t nchunk = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t c) -> int {
    const size_t off = c * kStageChunk, n = std::min(kStageChunk, bytes - off);
    HIPCHK(hipMemcpyAsync(P->stage_buf[c & 1], (const char*)src + off, n, hipMemcpyDeviceToHost,
                          P->side));
    HIPCHK(hipEventRecord(P->stage_ev[c & 1], P->side));
    return 0;
  };
  if (int rc = issue(0)) return rc;
  for (size_t c = 0; c < nchunk; c++) {
    // buffer (c+1)&1 held chunk c-1, whose host copy ended last iteration
    if (c + 1 < nchunk)
      if (int rc = issue(c + 1)) return rc;
    HIPCHK(hipEventSynchronize(P->stage_ev[c & 1]));
    const size_t off = c * kStageChunk, n = std::min(kStageChunk, bytes - off);
    g_copy.copy((char*)dst + off, P->stage_buf[c & 1], n);
  }
  return 0;
}

// True when a host call with these inputs may reuse the root search of the
// previous host call (same inputs, same launch configuration, and no other
// call has touched the pool since).
bool host_key_match(const Pool* P, const dazi_dk_desc* d, const float* vel, const float* vp,
                    const float* rho, const float* depz, const double* t) {
  const auto& k = P->hkey;
  const size_t nv = (size_t)d->nx * d->ny * d->nz;
  const int mode = d->vp_rho_mode == 1 ? 1 : 0;
  if (!k.valid || k.gen != P->gen || k.nx != d->nx || k.ny != d->ny || k.nz != d->nz ||
      k.kmax != d->kmax || k.mode != mode || k.minthk != d->minthk0 ||
      std::memcmp(&k.tune, &g_tune, sizeof g_tune) != 0)
    return false;
  if (std::memcmp(k.vel.data(), vel, nv * 4) || std::memcmp(k.depz.data(), depz, d->nz * 4) ||
      std::memcmp(k.t.data(), t, (size_t)d->kmax * 8))
    return false;
  if (mode && (std::memcmp(k.vp.data(), vp, nv * 4) || std::memcmp(k.rho.data(), rho, nv * 4)))
    return false;
  return true;
}

void host_key_store(Pool* P, const dazi_dk_desc* d, const float* vel, const float* vp,
                    const float* rho, const float* depz, const double* t) {
  auto& k = P->hkey;
  const size_t nv = (size_t)d->nx * d->ny * d->nz;
  k.valid = true;
  k.gen = P->gen;
  k.nx = d->nx;
  k.ny = d->ny;
  k.nz = d->nz;
  k.kmax = d->kmax;
  k.mode = d->vp_rho_mode == 1 ? 1 : 0;
  k.minthk = d->minthk0;
  k.tune = g_tune;
  k.vel.assign(vel, vel + nv);
  k.depz.assign(depz, depz + d->nz);
  k.t.assign(t, t + d->kmax);
  if (k.mode) {
    k.vp.assign(vp, vp + nv);
    k.rho.assign(rho, rho + nv);
  } else {
    k.vp.clear();
    k.rho.clear();
  }
}

// The whole stage on device pointers: root search, optional isotropic node
// kernels (svs/svp/srho) and optional TI splice (lsen).
int stage_impl(const dazi_dk_desc* d, const dazi_dk_tune* tune, const float* vel,
               const float* vp, const float* rho, const float* depz_dev,
               const float* depz_host, const double* tRc, double* pv, double* svs,
               double* svp, double* srho, float* lsen, int32_t* status, hipStream_t st,
               float* dcdA = nullptr, float* dcdL = nullptr, bool reuse = false) {
  dazi_dk_tune T;
  if (int rc = check_tune(tune, &T)) return rc;
  // plain mode's tile: the para.in control (pdsurf_tile_columns /
  // ti_tile_columns, CalSurfG.f90:173, depthkernelTI.f90:250) unless the
  // caller's own tune sets one
  if (d->tile_columns > 0 && !(tune && tune->tile_columns > 0)) T.tile_columns = d->tile_columns;
  if (d->iwave != 2) return set_err(DAZI_DK_EINVAL, "iwave=%d: only Rayleigh (2) is supported", d->iwave);
  if (d->igr != 0) return set_err(DAZI_DK_EINVAL, "igr=%d: only phase velocity (0) is supported", d->igr);
  if (d->nx < 1 || d->ny < 1 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d", d->nx, d->ny, d->nz,
                   d->kmax);
  if (!(d->minthk0 > 0.0f)) return set_err(DAZI_DK_EINVAL, "minthk must be > 0");
  if (d->vp_rho_mode == 1 && (!vp || !rho))
    return set_err(DAZI_DK_EINVAL, "vp_rho_mode=1 needs vp and rho");
  bool kern = svs || svp || srho;
  if (kern && !(svs && svp && srho))
    return set_err(DAZI_DK_EINVAL, "sen_vs, sen_vp, sen_rho must be all set or all NULL");
  if ((dcdA == nullptr) != (dcdL == nullptr))
    return set_err(DAZI_DK_EINVAL, "dc_dA and dc_dL must be both set or both NULL");
  const bool ti = lsen || dcdA;
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  Geometry* G;
  if (int rc = geometry(P, d, depz_dev, depz_host, st, &G)) return rc;
  const int L = G->L;
  void* nsubd = G->nsub;
  void* geo = G->recs;
  const float* geo_depz = (const float*)((const char*)geo + (size_t)L * dk::PREP_GEOM_BYTES);
  int ncol = d->nx * d->ny;
  int64_t nsolves = (int64_t)ncol * d->kmax;
  void *layers, *btp, *cc, *betmx, *bad, *croot, *stat, *work;
  if (int rc = pool_get(P, S_LAYERS, (size_t)ncol * L * 32, &layers)) return rc;
  if (int rc = pool_get(P, S_BTP, (size_t)ncol * L * 16, &btp)) return rc;
  double* btr = (double*)btp + (size_t)ncol * L;
  if (int rc = pool_get(P, S_CC, (size_t)ncol * 8, &cc)) return rc;
  if (int rc = pool_get(P, S_BETMX, (size_t)ncol * 4, &betmx)) return rc;
  if (int rc = pool_get(P, S_BAD, (size_t)ncol * 4, &bad)) return rc;
  if (int rc = pool_get(P, S_CROOT, (size_t)nsolves * 8, &croot)) return rc;
  if (status) {
    stat = status;
  } else if (int rc = pool_get(P, S_STATUS, (size_t)ncol * 4, &stat)) {
    return rc;
  }
  if (int rc = pool_get(P, S_WORK, 64, &work)) return rc;

  // every buffer is sized before the first launch, so no error return can
  // leave side-stream work unjoined
  void* tape = nullptr;
  int kern_nthr = 0;
  if (kern) {
    kern_nthr = kernel_threads(P, nsolves);
    if (int rc = pool_get(P, S_TAPE, dk::kern_tape_bytes(L, kern_nthr), &tape))
      return rc;
  }
  // the TI model prep needs only the inputs and the shared layer records: it
  // runs on the side stream from the start of the call
  void *til = nullptr, *tiscr = nullptr, *tiq = nullptr;
  int ti_nthr = 0;
  if (ti) {
    if (nsolves >= INT32_MAX)
      return set_err(DAZI_DK_ELIMIT, "TI splice: %lld solves (limit 2^31 - 1)", (long long)nsolves);
    if (int rc = pool_get(P, S_TI_LAYERS, (size_t)ncol * L * sizeof(dk::TiLayerDev), &til))
      return rc;
    ti_nthr = ti_threads(P, nsolves);
    if (int rc = pool_get(P, S_TI_SCRATCH, (size_t)dk::TI_SCRATCH * L * ti_nthr * 8, &tiscr))
      return rc;
    if (int rc = pool_get(P, S_TI_QUEUE, (size_t)(1 + nsolves) * 4, &tiq)) return rc;
  }
  P->have_events = g_timing;
  P->last_L = L;
  P->last_ti_threads = ti ? ti_nthr : 0;
  if (!reuse) P->gen++;
  rec(P, 0, st);
  if (!G->built) {
    HIPCHK(dk::launch_prep_geom(d->nz, depz_dev, d->minthk0, L, geo, (float*)geo_depz, st));
    G->built = true;
  }
  // a reused call keeps the flattened layers, start velocities and roots that
  // the previous host call left in these slots
  if (!reuse)
    HIPCHK(dk::launch_prep_nodes(ncol, d->nz, vel, d->vp_rho_mode == 1 ? vp : nullptr,
                                 d->vp_rho_mode == 1 ? rho : nullptr, depz_dev, geo_depz, L, geo,
                                 (double4*)layers, (double*)btp, btr, (double*)cc, (float*)betmx,
                                 (int*)bad, st));
  // the TI model prep on the same stream: a side stream joined by events
  // left the device idle for about 90 us between calls
  if (ti)
    HIPCHK(dk::launch_prep_ti(ncol, vel, d->vp_rho_mode == 1 ? vp : nullptr,
                              d->vp_rho_mode == 1 ? rho : nullptr, L, geo,
                              (dk::TiLayerDev*)til, (int*)tiq, st));
  rec(P, 1, st);
  dk::BatchDev B = {};
  B.ncol = ncol;
  B.uniform = 1;
  B.Lu = L;
  B.Ku = d->kmax;
  B.layers = (const double4*)layers;
  B.btp = (const double*)btp;
  B.btr = btr;
  B.periods = tRc;
  B.cc = (const double*)cc;
  B.betmx = (const float*)betmx;
  B.bad = (const int*)bad;
  B.croot = (double*)croot;
  B.pv = pv;
  B.status = (int*)stat;
  B.work = (int*)work;
  B.tol = T.refine_tol;
  if (!reuse) {
    if (int rc = run_forward(P, B, T, L, st)) return rc;
  } else {
    P->nforward = 0;
  }
  rec(P, 2, st);
  if (kern) {
    dk::KernDev O = {};
    O.node_mode = 1;
    O.nz = d->nz;
    O.nsub = (const int*)nsubd;
    O.nsolves = nsolves;
    O.dvs = svs;
    O.dvp = svp;
    O.drho = srho;
    HIPCHK(dk::launch_kernels(B, O, (double*)tape, dk::kern_checkpoints(L), kern_nthr, st));
  }
  rec(P, 3, st);
  if (ti) {
    HIPCHK(dk::launch_ti(ncol, d->kmax, L, d->nz, (const dk::TiLayerDev*)til, tRc, pv,
                         (const int*)nsubd, lsen, dcdA, dcdL, (double*)tiscr, ti_nthr, (int*)tiq,
                         (d->flags & DAZI_DK_TI_FAILED_NAN) != 0, st));
  }
  rec(P, 4, st);
  return 0;
}

int check_grow_desc(const dazi_dk_grow_desc* d) {
  if (!d) return set_err(DAZI_DK_EINVAL, "NULL descriptor");
  if (d->nvx < 1 || d->nvz < 1 || d->nz < 2 || d->kmax < 1 || d->nray < 0)
    return set_err(DAZI_DK_EINVAL, "bad sizes nvx=%d nvz=%d nz=%d kmax=%d nray=%lld", d->nvx,
                   d->nvz, d->nz, d->kmax, (long long)d->nray);
  if (!(d->ftol >= 0.0f)) return set_err(DAZI_DK_EINVAL, "ftol must be >= 0");
  const int64_t nmodel = 3 * (int64_t)d->nvx * d->nvz * (d->nz - 1);
  if (nmodel > INT32_MAX)
    return set_err(DAZI_DK_ELIMIT, "3*nparpi = %lld model columns exceed int32",
                   (long long)nmodel);
  if (d->nray > INT32_MAX || (int64_t)d->row_base + d->nray - 1 > INT32_MAX)
    return set_err(DAZI_DK_ELIMIT, "data rows beyond int32");
  return 0;
}

int grows_impl(const dazi_dk_grow_desc* d, const int32_t* ray_k, const int64_t* cell_off,
               const int32_t* cell, const float* fdm, const float* fdmc, const float* fdms,
               const float* vel, const double* svs, const double* svp, const double* srho,
               const float* lsen, int64_t cap, float* rw, int32_t* iw, int32_t* col,
               int64_t* row_nnz, int64_t* nar, hipStream_t st) {
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  void *cnt = row_nnz, *off;
  if (!cnt)
    if (int rc = pool_get(P, S_G_CNT, (size_t)(d->nray > 0 ? d->nray : 1) * 8, &cnt)) return rc;
  if (int rc = pool_get(P, S_G_ROWOFF, (size_t)(d->nray > 0 ? d->nray : 1) * 8, &off)) return rc;
  void *eff, *lst, *tsum;
  const size_t neff = (size_t)(d->nvx + 2) * (d->nvz + 2) * d->kmax * (d->nz - 1);
  if (int rc = pool_get(P, S_G_EFF, neff * 8, &eff)) return rc;
  if (int rc = pool_get(P, S_G_LSENT, neff * 4, &lst)) return rc;
  const long nt = dk::grow_scan_tiles(d->nray);
  if (int rc = pool_get(P, S_G_TSUM, (size_t)(nt > 0 ? nt : 1) * 8, &tsum)) return rc;
  dk::GrowDev g;
  g.nvx = d->nvx;
  g.nvz = d->nvz;
  g.nz = d->nz;
  g.K = d->kmax;
  g.nray = d->nray;
  g.row_base = d->row_base;
  g.ftol = d->ftol;
  g.ray_k = ray_k;
  g.cell_off = cell_off;
  g.cell = cell;
  g.fdm = fdm;
  g.fdmc = fdmc;
  g.fdms = fdms;
  g.vel = vel;
  g.svs = svs;
  g.svp = svp;
  g.srho = srho;
  g.lsen = lsen;
  g.effT = (double*)eff;
  g.lsenT = (float*)lst;
  g.tsum = (int64_t*)tsum;
  g.cap = cap;
  g.rw = rw;
  g.iw = iw;
  g.col = col;
  g.row_nnz = (int64_t*)cnt;
  g.row_off = (int64_t*)off;
  g.nar = nar;
  HIPCHK(dk::launch_grows(g, st));
  return 0;
}

// The host entry points (the Fortran drop-in): inputs copied to the pool's
// host-path slots, the stage, outputs copied back through pinned staging.  A
// call whose inputs are byte-identical to the previous host call's (the G
// builder calls depthkernelTI and then depthkernel on the same model,
// CalSurfGAniso_Joint.f90:478,486) reuses that call's root search and runs
// only its own kernel pass.
int host_stage(const dazi_dk_desc* d, const float* vel, const float* vp, const float* rho,
               const float* depz, const double* tRc, double* pvRc, double* svs, double* svp,
               double* srho, float* lsen) {
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  const size_t ncol = (size_t)d->nx * d->ny, nz = d->nz, K = d->kmax;
  const bool mode1 = d->vp_rho_mode == 1;
  const bool kern = svs != nullptr;
  void *dvel, *dvp = nullptr, *drho = nullptr, *ddepz, *dt, *dpv, *ds1 = nullptr, *ds2 = nullptr,
              *ds3 = nullptr, *dls = nullptr;
  if (int rc = pool_get(P, S_H_VEL, ncol * nz * 4, &dvel)) return rc;
  if (mode1) {
    if (int rc = pool_get(P, S_H_VP, ncol * nz * 4, &dvp)) return rc;
    if (int rc = pool_get(P, S_H_RHO, ncol * nz * 4, &drho)) return rc;
  }
  if (int rc = pool_get(P, S_H_DEPZ, nz * 4, &ddepz)) return rc;
  if (int rc = pool_get(P, S_H_T, K * 8, &dt)) return rc;
  if (int rc = pool_get(P, S_H_PV, ncol * K * 8, &dpv)) return rc;
  if (kern) {
    if (int rc = pool_get(P, S_H_SVS, ncol * K * nz * 8, &ds1)) return rc;
    if (int rc = pool_get(P, S_H_SVP, ncol * K * nz * 8, &ds2)) return rc;
    if (int rc = pool_get(P, S_H_SRHO, ncol * K * nz * 8, &ds3)) return rc;
  }
  if (lsen)
    if (int rc = pool_get(P, S_H_LSEN, ncol * K * (nz - 1) * 4, &dls)) return rc;
  // (sizing can grow a slot and so retire the previous call's buffers: the
  // generation check below runs after it)
  const bool reuse = host_key_match(P, d, vel, vp, rho, depz, tRc);
  hipStream_t st = nullptr;
  if (!reuse) {
    HIPCHK(hipMemcpy(dvel, vel, ncol * nz * 4, hipMemcpyHostToDevice));
    if (mode1) {
      HIPCHK(hipMemcpy(dvp, vp, ncol * nz * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(drho, rho, ncol * nz * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ddepz, depz, nz * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dt, tRc, K * 8, hipMemcpyHostToDevice));
  }
  int rc = stage_impl(d, nullptr, (const float*)dvel, (const float*)dvp, (const float*)drho,
                      (const float*)ddepz, depz, (const double*)dt, (double*)dpv, (double*)ds1,
                      (double*)ds2, (double*)ds3, (float*)dls, nullptr, st, nullptr, nullptr,
                      reuse);
  if (rc) {
    P->hkey.valid = false;
    return rc;
  }
  host_key_store(P, d, vel, vp, rho, depz, tRc);
  if (int r = copy_out(P, pvRc, dpv, ncol * K * 8, st)) return r;
  if (kern) {
    if (int r = copy_out(P, svs, ds1, ncol * K * nz * 8, st)) return r;
    if (int r = copy_out(P, svp, ds2, ncol * K * nz * 8, st)) return r;
    if (int r = copy_out(P, srho, ds3, ncol * K * nz * 8, st)) return r;
  }
  if (lsen)
    if (int r = copy_out(P, lsen, dls, ncol * K * (nz - 1) * 4, st)) return r;
  return 0;
}

}  // namespace

extern "C" {

int32_t dazi_dk_abi_version(void) { return DAZI_DK_ABI_VERSION; }

const char* dazi_dk_last_error(void) { return g_err.c_str(); }

int32_t dazi_dk_set_tune(const dazi_dk_tune* tune) {
  dazi_dk_tune T;
  dazi_dk_tune save = g_tune;
  if (!tune) return set_err(DAZI_DK_EINVAL, "tune is NULL");
  if (int rc = check_tune(tune, &T)) {
    g_tune = save;
    return rc;
  }
  g_tune = T;
  return 0;
}

int32_t dazi_dk_get_tune(dazi_dk_tune* tune) {
  if (!tune) return set_err(DAZI_DK_EINVAL, "tune is NULL");
  *tune = g_tune;
  return 0;
}

int32_t dazi_dk_enable_timing(int32_t on) {
  g_timing = on ? 1 : 0;
  return 0;
}

int32_t dazi_dk_last_timings(double* out, int32_t n) {
  Pool* P;
  if (int rc = get_pool(&P)) return -rc;
  if (!P->have_events || n < 1) return 0;
  double v[7] = {0, 0, 0, 0, (double)P->nforward, (double)P->last_L,
                 (double)P->last_ti_threads};
  for (int i = 0; i < 4; i++) {
    float ms = 0.f;
    if (hipEventSynchronize(P->ev[i + 1]) != hipSuccess) return 0;
    if (hipEventElapsedTime(&ms, P->ev[i], P->ev[i + 1]) != hipSuccess) return 0;
    v[i] = ms;
  }
  int m = n < 7 ? n : 7;
  for (int i = 0; i < m; i++) out[i] = v[i];
  return m;
}

int32_t dazi_dk_effective_dev(const dazi_dk_desc* d, const float* vel, const double* sen_vs,
                              const double* sen_vp, const double* sen_rho, const double* grad_c,
                              double* eff, double* grad_vs, void* stream) {
  if (!d || !vel || !sen_vs || !sen_vp || !sen_rho)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (grad_vs && !grad_c) return set_err(DAZI_DK_EINVAL, "grad_vs needs grad_c");
  if (d->nx < 1 || d->ny < 1 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d", d->nx, d->ny, d->nz,
                   d->kmax);
  if (!eff && !grad_vs) return 0;
  HIPCHK(dk::launch_effective(d->nx * d->ny, d->nz, d->kmax, vel, sen_vs, sen_vp, sen_rho, grad_c,
                              eff, grad_vs, (hipStream_t)stream));
  return 0;
}

int32_t dazi_dk_cell_sum_dev(const double* gl, int64_t ld, int32_t nl, int64_t c0, int64_t n,
                             int32_t ncol_real, int32_t nreal, const int64_t* cell,
                             int32_t ncells, double* acc, void* stream) {
  if (nl < 1 || ncol_real < 1 || nreal < 1 || ncells < 1 || c0 < 0 || n < 0 || ld < n)
    return set_err(DAZI_DK_EINVAL, "bad sizes nl=%d ncol_real=%d nreal=%d ncells=%d c0=%lld n=%lld "
                   "ld=%lld", nl, ncol_real, nreal, ncells, (long long)c0, (long long)n,
                   (long long)ld);
  if (n == 0) return 0;
  if (!gl || !cell || !acc) return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if ((c0 + n - 1) / ncol_real >= nreal)
    return set_err(DAZI_DK_EINVAL, "columns %lld..%lld reach beyond %d realizations of %d columns",
                   (long long)c0, (long long)(c0 + n - 1), nreal, ncol_real);
  const int64_t nr = (c0 + n - 1) / ncol_real - c0 / ncol_real + 1;
  if (nr * nl > INT32_MAX) return set_err(DAZI_DK_ELIMIT, "too many (realization, layer) pairs");
  HIPCHK(dk::launch_cell_sum(gl, ld, nl, c0, n, ncol_real, cell, ncells, acc,
                             (hipStream_t)stream));
  return 0;
}

int32_t dazi_dk_grows_dev(const dazi_dk_grow_desc* d, const int32_t* ray_k,
                          const int64_t* cell_off, const int32_t* cell, const float* fdm,
                          const float* fdmc, const float* fdms, const float* vel,
                          const double* sen_vs, const double* sen_vp, const double* sen_rho,
                          const float* lsen, int64_t cap, float* rw, int32_t* iw, int32_t* col,
                          int64_t* row_nnz, int64_t* nar, void* stream) {
  if (int rc = check_grow_desc(d)) return rc;
  if (!nar) return set_err(DAZI_DK_EINVAL, "NULL nar");
  if (d->nray > 0 && (!ray_k || !cell_off || !cell || !fdm || !fdmc || !fdms || !vel ||
                      !sen_vs || !sen_vp || !sen_rho || !lsen))
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (cap < 0 || (cap > 0 && (!rw || !iw || !col)))
    return set_err(DAZI_DK_EINVAL, "cap > 0 needs rw, iw and col");
  return grows_impl(d, ray_k, cell_off, cell, fdm, fdmc, fdms, vel, sen_vs, sen_vp, sen_rho, lsen,
                    cap, rw, iw, col, row_nnz, nar, (hipStream_t)stream);
}

int32_t dazi_dk_grows(const dazi_dk_grow_desc* d, const int32_t* ray_k, const int64_t* cell_off,
                      const int32_t* cell, const float* fdm, const float* fdmc, const float* fdms,
                      const float* vel, const double* sen_vs, const double* sen_vp,
                      const double* sen_rho, const float* lsen, int64_t cap, float* rw,
                      int32_t* iw, int32_t* col, int64_t* nar) {
  if (int rc = check_grow_desc(d)) return rc;
  if (!nar || !cell_off) return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (d->nray > 0 && (!ray_k || !cell || !fdm || !fdmc || !fdms || !vel || !sen_vs || !sen_vp ||
                      !sen_rho || !lsen))
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  const int64_t nray = d->nray;
  const int cells = d->nvx * d->nvz;
  if (cell_off[0] != 0) return set_err(DAZI_DK_EINVAL, "cell_off[0] must be 0");
  for (int64_t r = 0; r < nray; r++) {
    if (cell_off[r + 1] < cell_off[r])
      return set_err(DAZI_DK_EINVAL, "cell_off decreases at ray %lld", (long long)r);
    if (ray_k[r] < 0 || ray_k[r] >= d->kmax)
      return set_err(DAZI_DK_EINVAL, "ray %lld: period index %d outside 0..%d", (long long)r,
                     ray_k[r], d->kmax - 1);
    for (int64_t i = cell_off[r]; i < cell_off[r + 1]; i++) {
      if (cell[i] < 0 || cell[i] >= cells)
        return set_err(DAZI_DK_EINVAL, "ray %lld: cell %d outside 0..%d", (long long)r, cell[i],
                       cells - 1);
      if (i > cell_off[r] && cell[i] <= cell[i - 1])
        return set_err(DAZI_DK_EINVAL, "ray %lld: cells not strictly increasing", (long long)r);
    }
  }
  if (cap < 0 || (cap > 0 && (!rw || !iw || !col)))
    return set_err(DAZI_DK_EINVAL, "cap > 0 needs rw, iw and col");
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  P->gen++;  // the host-path input slots are reused below
  const size_t ncol = (size_t)(d->nvx + 2) * (d->nvz + 2), nz = d->nz, K = d->kmax;
  const size_t nc = (size_t)cell_off[nray], nr = (size_t)(nray > 0 ? nray : 1);
  void *dk_, *doff, *dcell, *dw, *dvel, *d1, *d2, *d3, *dls, *dnar, *drw = nullptr,
       *diw = nullptr, *dcol = nullptr;
  if (int rc = pool_get(P, S_G_RAYK, nr * 4, &dk_)) return rc;
  if (int rc = pool_get(P, S_G_CELLOFF, (nr + 1) * 8, &doff)) return rc;
  if (int rc = pool_get(P, S_G_CELL, (nc > 0 ? nc : 1) * 4, &dcell)) return rc;
  if (int rc = pool_get(P, S_G_W, (nc > 0 ? nc : 1) * 12, &dw)) return rc;
  if (int rc = pool_get(P, S_H_VEL, ncol * nz * 4, &dvel)) return rc;
  if (int rc = pool_get(P, S_H_SVS, ncol * K * nz * 8, &d1)) return rc;
  if (int rc = pool_get(P, S_H_SVP, ncol * K * nz * 8, &d2)) return rc;
  if (int rc = pool_get(P, S_H_SRHO, ncol * K * nz * 8, &d3)) return rc;
  if (int rc = pool_get(P, S_H_LSEN, ncol * K * (nz - 1) * 4, &dls)) return rc;
  if (int rc = pool_get(P, S_G_NAR, 8, &dnar)) return rc;
  if (cap > 0) {
    if (int rc = pool_get(P, S_G_RW, (size_t)cap * 4, &drw)) return rc;
    if (int rc = pool_get(P, S_G_IW, (size_t)cap * 4, &diw)) return rc;
    if (int rc = pool_get(P, S_G_COL, (size_t)cap * 4, &dcol)) return rc;
  }
  float* w = (float*)dw;
  if (nray > 0) {
    HIPCHK(hipMemcpy(dk_, ray_k, nray * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(doff, cell_off, (nray + 1) * 8, hipMemcpyHostToDevice));
    if (nc > 0) {
      HIPCHK(hipMemcpy(dcell, cell, nc * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(w, fdm, nc * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(w + nc, fdmc, nc * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(w + 2 * nc, fdms, nc * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(dvel, vel, ncol * nz * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d1, sen_vs, ncol * K * nz * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d2, sen_vp, ncol * K * nz * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d3, sen_rho, ncol * K * nz * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dls, lsen, ncol * K * (nz - 1) * 4, hipMemcpyHostToDevice));
  }
  hipStream_t st = nullptr;
  if (int rc = grows_impl(d, (const int32_t*)dk_, (const int64_t*)doff, (const int32_t*)dcell, w,
                          w + nc, w + 2 * nc, (const float*)dvel, (const double*)d1,
                          (const double*)d2, (const double*)d3, (const float*)dls, cap,
                          (float*)drw, (int32_t*)diw, (int32_t*)dcol, nullptr, (int64_t*)dnar, st))
    return rc;
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(nar, dnar, 8, hipMemcpyDeviceToHost));
  const int64_t n = *nar < cap ? *nar : cap;
  if (n > 0) {
    HIPCHK(hipMemcpy(rw, drw, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(iw, diw, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(col, dcol, (size_t)n * 4, hipMemcpyDeviceToHost));
  }
  if (*nar > cap)
    return set_err(DAZI_DK_ELIMIT, "%lld entries exceed cap %lld (nar holds the count)",
                   (long long)*nar, (long long)cap);
  return 0;
}

int32_t dazi_dk_azimuth_map_dev(const dazi_dk_azim_desc* d, const float* lsen, const float* gcf,
                                const float* gsf, const double* tRcV, float* cos_rc,
                                float* sin_rc, float* amp, float* amp_rela, float* angle,
                                void* stream) {
  if (!d) return set_err(DAZI_DK_EINVAL, "NULL descriptor");
  if (d->nx < 3 || d->ny < 3 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d (nx, ny >= 3)", d->nx,
                   d->ny, d->nz, d->kmax);
  if (!lsen || !gcf || !gsf || !cos_rc || !sin_rc || (amp_rela && !tRcV))
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  HIPCHK(dk::launch_azimuth_map(d->nx - 2, d->ny - 2, d->nz, d->kmax, lsen, gcf, gsf, tRcV,
                                cos_rc, sin_rc, amp, amp_rela, angle, (hipStream_t)stream));
  return 0;
}

int32_t dazi_dk_azimuth_map(const dazi_dk_azim_desc* d, const float* lsen, const float* gcf,
                            const float* gsf, const double* tRcV, float* cos_rc, float* sin_rc,
                            float* amp, float* amp_rela, float* angle) {
  if (!d) return set_err(DAZI_DK_EINVAL, "NULL descriptor");
  if (d->nx < 3 || d->ny < 3 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d (nx, ny >= 3)", d->nx,
                   d->ny, d->nz, d->kmax);
  if (!lsen || !gcf || !gsf || !tRcV || !cos_rc || !sin_rc || !amp || !amp_rela || !angle)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  P->gen++;  // the host-path output slot S_H_LSEN is reused below
  const size_t ncol = (size_t)d->nx * d->ny, nd = d->nz - 1, K = d->kmax;
  const size_t cells = (size_t)(d->nx - 2) * (d->ny - 2);
  void *dl, *dg, *dt, *dout;
  if (int rc = pool_get(P, S_H_LSEN, ncol * K * nd * 4, &dl)) return rc;
  if (int rc = pool_get(P, S_A_G, cells * nd * 8, &dg)) return rc;
  if (int rc = pool_get(P, S_A_T, cells * K * 8, &dt)) return rc;
  if (int rc = pool_get(P, S_A_OUT, cells * K * 4 * 5, &dout)) return rc;
  float* g = (float*)dg;
  float* o = (float*)dout;
  const size_t m = cells * K;
  HIPCHK(hipMemcpy(dl, lsen, ncol * K * nd * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g, gcf, cells * nd * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g + cells * nd, gsf, cells * nd * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dt, tRcV, m * 8, hipMemcpyHostToDevice));
  HIPCHK(dk::launch_azimuth_map(d->nx - 2, d->ny - 2, d->nz, d->kmax, (const float*)dl, g,
                                g + cells * nd, (const double*)dt, o, o + m, o + 2 * m, o + 3 * m,
                                o + 4 * m, nullptr));
  HIPCHK(hipStreamSynchronize(nullptr));
  float* outs[5] = {cos_rc, sin_rc, amp, amp_rela, angle};
  for (int q = 0; q < 5; q++) HIPCHK(hipMemcpy(outs[q], o + q * m, m * 4, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

// ---- first-arrival traveltimes ----

struct dazi_dk_travel_plan {
  int dev = -1;
  dk::TravelDev p = {};
  std::vector<void*> bufs;
  ~dazi_dk_travel_plan() {
    for (void* b : bufs) (void)hipFree(b);
  }
};

namespace {

template <class T>
int dev_alloc(dazi_dk_travel_plan* P, size_t n, T** out) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 8)));
  P->bufs.push_back(q);
  *out = (T*)q;
  return 0;
}

template <class T>
int dev_upload(dazi_dk_travel_plan* P, const T* h, size_t n, const T** out) {
  T* d;
  if (int rc = dev_alloc(P, n, &d)) return rc;
  if (n) HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  *out = d;
  return 0;
}

// Grid constants of CalSurfGAniso_Joint.f90:396-409 and the refined spacing
// of :570-571, in the reference's single-precision operation order.
void travel_grid(const dazi_dk_travel_desc* d, dk::TravelDev& p) {
#pragma clang fp contract(off)
  const float pi = 3.1415926535898f;
  const float dvx = d->dvxd * pi / 180.0f, dvz = d->dvzd * pi / 180.0f;
  p.gox = (90.0f - d->goxd) * pi / 180.0f;
  p.goz = d->gozd * pi / 180.0f;
  p.dnx = dvx / (float)d->gdx;
  p.dnz = dvz / (float)d->gdz;
  p.drnx = dvx / (float)(d->gdx * d->sgdl);
  p.drnz = dvz / (float)(d->gdz * d->sgdl);
  p.nnx = (d->nvx - 1) * d->gdx + 1;
  p.nnz = (d->nvz - 1) * d->gdz + 1;
  p.maxbt_c = (int)std::lround(d->snb * (float)p.nnx * (float)p.nnz);
}

// Cell index of a coordinate, INT((x - o)/d) + 1 as the reference computes it.
int cell_index(float x, float o, float dn) {
#pragma clang fp contract(off)
  return (int)((x - o) / dn) + 1;
}

// The sin tables the marching reads (libm sinf, as the reference):
// risti = earth*sin(colatitude) per column of the coarse grid and of each
// refined grid (by its first coarse column vnl), and srtimes' dpl.
void travel_tables(const dazi_dk_travel_desc* d, dk::TravelDev& p, std::vector<float>& rc,
                   std::vector<float>& rr) {
#pragma clang fp contract(off)
  rc.resize(p.nnx);
  for (int ix = 1; ix <= p.nnx; ix++) rc[ix - 1] = d->earth * sinf(p.gox + (float)(ix - 1) * p.dnx);
  rr.assign((size_t)p.nnx * p.nrx_max, 0.0f);
  for (int vnl = 1; vnl <= p.nnx; vnl++) {
    const float gorx = p.gox + p.dnx * (float)(vnl - 1);
    for (int ix = 1; ix <= p.nrx_max; ix++)
      rr[(size_t)(vnl - 1) * p.nrx_max + ix - 1] = d->earth * sinf(gorx + (float)(ix - 1) * p.drnx);
  }
  float dpl = p.dnx * d->earth;
  float rd1 = p.dnz * d->earth * sinf(p.gox);
  if (rd1 < dpl) dpl = rd1;
  rd1 = p.dnz * d->earth * sinf(p.gox + (float)(p.nnx - 1) * p.dnx);
  if (rd1 < dpl) dpl = rd1;
  p.dpl = dpl;
}

}  // namespace

extern "C" {

int32_t dazi_dk_travel_plan_create(const dazi_dk_travel_desc* d, int64_t njob,
                                   const int32_t* job_k, const float* scx, const float* scz,
                                   const int64_t* rec_off, const float* rcx, const float* rcz,
                                   dazi_dk_travel_plan** out) {
  if (!d || !out) return set_err(DAZI_DK_EINVAL, "NULL descriptor or plan");
  *out = nullptr;
  if (d->nvx < 2 || d->nvz < 2 || d->gdx < 1 || d->gdz < 1 || d->sgdl < 1 || d->sgs < 1 ||
      (d->asgr != 0 && d->asgr != 1) || (d->fom != 0 && d->fom != 1) || !(d->snb > 0.0f) ||
      !(d->earth > 0.0f) || d->kmax < 1 || d->slots < 0 || !(d->dvxd > 0.0f) ||
      !(d->dvzd > 0.0f))
    return set_err(DAZI_DK_EINVAL,
                   "bad travel descriptor (nvx=%d nvz=%d gdx=%d gdz=%d sgdl=%d sgs=%d asgr=%d "
                   "fom=%d snb=%g earth=%g kmax=%d slots=%d)",
                   d->nvx, d->nvz, d->gdx, d->gdz, d->sgdl, d->sgs, d->asgr, d->fom, d->snb,
                   d->earth, d->kmax, d->slots);
  if (njob < 0 || (njob > 0 && (!job_k || !scx || !scz || !rec_off)))
    return set_err(DAZI_DK_EINVAL, "bad job arrays (njob=%lld)", (long long)njob);
  auto P = std::make_unique<dazi_dk_travel_plan>();
  dk::TravelDev& p = P->p;
  p.nvx = d->nvx;
  p.nvz = d->nvz;
  p.gdx = d->gdx;
  p.gdz = d->gdz;
  p.sgdl = d->sgdl;
  p.sgs = d->sgs;
  p.asgr = d->asgr;
  p.fom = d->fom;
  p.K = d->kmax;
  p.ncol = (long)(d->nvx + 2) * (d->nvz + 2);
  p.earth = d->earth;
  p.snb = d->snb;
  travel_grid(d, p);
  if (p.nnx > 32767 || p.nnz > 32767)
    return set_err(DAZI_DK_EINVAL, "propagation grid %d x %d exceeds 32767 nodes per side", p.nnz,
                   p.nnx);
  p.nrx_max = std::min(2 * d->sgs, p.nnx - 1) * d->sgdl + 1;
  const int nrz_max = std::min(2 * d->sgs, p.nnz - 1) * d->sgdl + 1;
  if (d->asgr && (p.nrx_max > 32767 || nrz_max > 32767))
    return set_err(DAZI_DK_EINVAL, "refined source grid exceeds 32767 nodes per side");
  p.nr_max = d->asgr ? p.nrx_max * nrz_max : 1;
  p.nrp_max = d->asgr ? (long)(p.nrx_max + 6) * (nrz_max + 6) : 1;
  const int64_t nrec = njob > 0 ? rec_off[njob] : 0;
  if (njob > 0 && rec_off[0] != 0) return set_err(DAZI_DK_EINVAL, "rec_off[0] must be 0");
  for (int64_t q = 0; q < njob; q++) {
    if (rec_off[q + 1] < rec_off[q])
      return set_err(DAZI_DK_EINVAL, "rec_off decreases at job %lld", (long long)q);
    if (job_k[q] < 0 || job_k[q] >= d->kmax)
      return set_err(DAZI_DK_EINVAL, "job %lld: period index %d outside 0..%d", (long long)q,
                     job_k[q], d->kmax - 1);
    const int isx = std::isfinite(scx[q]) ? cell_index(scx[q], p.gox, p.dnx) : 0;
    const int isz = std::isfinite(scz[q]) ? cell_index(scz[q], p.goz, p.dnz) : 0;
    if (!std::isfinite(scx[q]) || !std::isfinite(scz[q]) || isx < 1 || isx > p.nnx || isz < 1 ||
        isz > p.nnz)
      return set_err(DAZI_DK_EINVAL, "job %lld: source (%g, %g) lies outside the model",
                     (long long)q, scx[q], scz[q]);
  }
  if (nrec > 0 && (!rcx || !rcz)) return set_err(DAZI_DK_EINVAL, "NULL receiver arrays");
  std::vector<float> srx((size_t)nrec);
  for (int64_t r = 0; r < nrec; r++) {
    const bool fin = std::isfinite(rcx[r]) && std::isfinite(rcz[r]);
    const int irx = fin ? cell_index(rcx[r], p.gox, p.dnx) : 0;
    const int irz = fin ? cell_index(rcz[r], p.goz, p.dnz) : 0;
    if (!fin || irx < 1 || irx > p.nnx || irz < 1 || irz > p.nnz)
      return set_err(DAZI_DK_EINVAL, "receiver %lld (%g, %g) lies outside the model",
                     (long long)r, rcx[r], rcz[r]);
    srx[r] = sinf(rcx[r]);
  }
  // tree capacity: the coarse grid's, or the larger refined grid's
  const int mx = std::max(p.nrx_max, p.nnx), mz = std::max(nrz_max, p.nnz);
  const long maxbt_max =
      std::max<long>(p.maxbt_c, d->asgr ? std::lround(d->snb * (float)mx * (float)mz) : 0);
  p.spill_cap = std::max<long>(maxbt_max - 1024, 0);
  HIPCHK(hipGetDevice(&P->dev));
  Pool* pool;
  if (int rc = get_pool(&pool)) return rc;
  const long nn = (long)p.nnx * p.nnz, ncp = (long)(p.nnx + 6) * (p.nnz + 6);
  const double per_slot = 8.0 * ncp + 8.0 * p.nrp_max + 4.0 * p.nr_max + 8.0 * p.spill_cap;
  long slots = d->slots > 0 ? d->slots : (long)pool->ncu * 16;
  slots = std::min<long>(slots, std::max<long>((long)(8.0e9 / per_slot), 1));
  slots = std::max<long>(std::min<long>(slots, njob), 1);
  p.slots = (int)slots;
  p.njob = njob;
  std::vector<float> rc_tab, rr_tab;
  p.nrx_max = std::max(p.nrx_max, 1);
  travel_tables(d, p, rc_tab, rr_tab);
  dazi_dk_travel_plan* Q = P.get();
  std::vector<int64_t> zero_off(1, 0);
  if (int rc = dev_upload(Q, job_k, (size_t)njob, &p.job_k)) return rc;
  if (int rc = dev_upload(Q, scx, (size_t)njob, &p.scx)) return rc;
  if (int rc = dev_upload(Q, scz, (size_t)njob, &p.scz)) return rc;
  if (int rc = dev_upload(Q, njob ? rec_off : zero_off.data(), (size_t)njob + 1, &p.rec_off))
    return rc;
  if (int rc = dev_upload(Q, rcx, (size_t)nrec, &p.rcx)) return rc;
  if (int rc = dev_upload(Q, rcz, (size_t)nrec, &p.rcz)) return rc;
  if (int rc = dev_upload(Q, srx.data(), srx.size(), &p.sin_rcx)) return rc;
  if (int rc = dev_upload(Q, rc_tab.data(), rc_tab.size(), &p.risti_c)) return rc;
  if (int rc = dev_upload(Q, rr_tab.data(), rr_tab.size(), &p.risti_r)) return rc;
  if (int rc = dev_alloc(Q, (size_t)nn * p.K, &p.veln)) return rc;
  if (int rc = dev_alloc(Q, (size_t)ncp * slots, &p.node_c)) return rc;
  if (int rc = dev_alloc(Q, (size_t)p.nrp_max * slots, &p.node_r)) return rc;
  if (int rc = dev_alloc(Q, (size_t)p.nr_max * slots, &p.vel_r)) return rc;
  if (int rc = dev_alloc(Q, (size_t)std::max<long>(p.spill_cap, 1) * slots, &p.spill)) return rc;
  *out = P.release();
  return 0;
}

int32_t dazi_dk_travel_run_dev(dazi_dk_travel_plan* P, const double* pvRc, float* times,
                               float* ttn, int32_t* status, void* stream) {
  if (!P) return set_err(DAZI_DK_EINVAL, "NULL plan");
  if (!pvRc || (P->p.njob > 0 && !status)) return set_err(DAZI_DK_EINVAL, "NULL pvRc or status");
  if (!times && P->p.njob > 0) return set_err(DAZI_DK_EINVAL, "NULL times");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != P->dev)
    return set_err(DAZI_DK_EINVAL, "plan belongs to device %d, current device is %d", P->dev, dev);
  dk::TravelDev p = P->p;
  p.pv = pvRc;
  p.times = times;
  p.ttn = ttn;
  p.status = status;
  HIPCHK(dk::launch_travel(p, (hipStream_t)stream));
  return 0;
}

int32_t dazi_dk_travel_plan_destroy(dazi_dk_travel_plan* P) {
  if (!P) return 0;
  int dev = -1;
  const int own = P->dev;
  HIPCHK(hipGetDevice(&dev));
  if (dev != own) HIPCHK(hipSetDevice(own));
  HIPCHK(hipDeviceSynchronize());
  delete P;
  if (dev != own) HIPCHK(hipSetDevice(dev));
  return 0;
}

int32_t dazi_dk_travel(const dazi_dk_travel_desc* d, int64_t njob, const int32_t* job_k,
                       const float* scx, const float* scz, const int64_t* rec_off,
                       const float* rcx, const float* rcz, const double* pvRc, float* times,
                       float* ttn, int32_t* status) {
  if (!pvRc || (njob > 0 && (!times || !status)))
    return set_err(DAZI_DK_EINVAL, "NULL pvRc, times or status");
  dazi_dk_travel_plan* P = nullptr;
  if (int rc = dazi_dk_travel_plan_create(d, njob, job_k, scx, scz, rec_off, rcx, rcz, &P))
    return rc;
  std::unique_ptr<dazi_dk_travel_plan> own(P);
  const dk::TravelDev& p = P->p;
  const size_t nn = (size_t)p.nnx * p.nnz, nrec = njob > 0 ? (size_t)rec_off[njob] : 0;
  double* dpv;
  float *dt, *dttn = nullptr;
  int32_t* dst;
  if (int rc = dev_alloc(P, (size_t)p.ncol * p.K, &dpv)) return rc;
  if (int rc = dev_alloc(P, nrec, &dt)) return rc;
  if (ttn)
    if (int rc = dev_alloc(P, nn * (size_t)njob, &dttn)) return rc;
  if (int rc = dev_alloc(P, (size_t)njob, &dst)) return rc;
  HIPCHK(hipMemcpy(dpv, pvRc, (size_t)p.ncol * p.K * 8, hipMemcpyHostToDevice));
  if (int rc = dazi_dk_travel_run_dev(P, dpv, dt, dttn, dst, nullptr)) return rc;
  HIPCHK(hipStreamSynchronize(nullptr));
  if (nrec) HIPCHK(hipMemcpy(times, dt, nrec * 4, hipMemcpyDeviceToHost));
  if (ttn && njob) HIPCHK(hipMemcpy(ttn, dttn, nn * njob * 4, hipMemcpyDeviceToHost));
  if (njob) HIPCHK(hipMemcpy(status, dst, (size_t)njob * 4, hipMemcpyDeviceToHost));
  return 0;
}

int32_t dazi_dk_pool_bytes(int64_t* out) {
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  int64_t n = (int64_t)(P->retired_bytes + P->geo_bytes);
  for (int s = 0; s < S_COUNT; s++) n += (int64_t)P->n[s];
  *out = n;
  return 0;
}

int32_t dazi_dk_node_layers(int32_t nz, const float* depz, float minthk, int32_t* L) {
  if (!depz || !L || nz < 2 || !(minthk > 0.0f))
    return set_err(DAZI_DK_EINVAL, "bad arguments (nz=%d, minthk=%g)", nz, minthk);
  std::vector<int> nsub;
  const int Lc = node_geometry(nz, depz, minthk, &nsub);
  if (Lc < 2) return set_err(DAZI_DK_EINVAL, "depz must be strictly increasing");
  *L = Lc;
  return 0;
}

int32_t dazi_dk_ti_kernels_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune,
                               const float* vel, const float* vp, const float* rho,
                               const float* depz, const double* tRc, double* pvRc,
                               float* Lsen_Gsc, float* dc_dA, float* dc_dL, int32_t* status,
                               void* stream) {
  if (!desc || !vel || !depz || !tRc || !pvRc || !dc_dA || !dc_dL)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  return stage_impl(desc, tune, vel, vp, rho, depz, nullptr, tRc, pvRc, nullptr, nullptr, nullptr,
                    Lsen_Gsc, status, (hipStream_t)stream, dc_dA, dc_dL);
}

int32_t dazi_dk_ti_vjp_dev(int32_t ncol, int32_t L, int32_t kmax, const float* dc_dA,
                           const float* dc_dL, const double* grad_c, double* grad_A,
                           double* grad_L, void* stream) {
  if (ncol < 0 || L < 1 || kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes ncol=%d L=%d kmax=%d", ncol, L, kmax);
  if (!dc_dA || !dc_dL || !grad_c || (!grad_A && !grad_L))
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (ncol == 0) return 0;
  HIPCHK(dk::launch_ti_vjp(ncol, L, kmax, dc_dA, dc_dL, grad_c, grad_A, grad_L,
                           (hipStream_t)stream));
  return 0;
}

int32_t dazi_dk_reset_geometry(void) {
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  P->cur = nullptr;
  return 0;
}

int32_t dazi_dk_iso_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                        const float* vp, const float* rho, const float* depz, const double* tRc,
                        double* pvRc, double* sen_vsRc, double* sen_vpRc, double* sen_rhoRc,
                        int32_t* status, void* stream) {
  if (!desc || !vel || !depz || !tRc || !pvRc)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  hipStream_t st = (hipStream_t)stream;
  return stage_impl(desc, tune, vel, vp, rho, depz, nullptr, tRc, pvRc, sen_vsRc, sen_vpRc, sen_rhoRc,
                    nullptr, status, st);
}

static int dev_entry(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                     const float* vp, const float* rho, const float* depz, const double* tRc,
                     double* pvRc, double* svs, double* svp, double* srho, float* lsen,
                     int32_t* status, void* stream) {
  if (!desc || !vel || !depz || !tRc || !pvRc)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  hipStream_t st = (hipStream_t)stream;
  return stage_impl(desc, tune, vel, vp, rho, depz, nullptr, tRc, pvRc, svs, svp, srho, lsen, status,
                    st);
}

int32_t dazi_dk_ti_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                       const float* vp, const float* rho, const float* depz, const double* tRc,
                       double* pvRc, float* Lsen_Gsc, int32_t* status, void* stream) {
  if (!Lsen_Gsc) return set_err(DAZI_DK_EINVAL, "NULL Lsen_Gsc");
  return dev_entry(desc, tune, vel, vp, rho, depz, tRc, pvRc, nullptr, nullptr, nullptr, Lsen_Gsc,
                   status, stream);
}

int32_t dazi_dk_stage_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                          const float* vp, const float* rho, const float* depz, const double* tRc,
                          double* pvRc, double* sen_vsRc, double* sen_vpRc, double* sen_rhoRc,
                          float* Lsen_Gsc, int32_t* status, void* stream) {
  return dev_entry(desc, tune, vel, vp, rho, depz, tRc, pvRc, sen_vsRc, sen_vpRc, sen_rhoRc,
                   Lsen_Gsc, status, stream);
}

int32_t dazi_dk_ti(const dazi_dk_desc* d, const float* vel, const float* depz, const double* tRc,
                   double* pvRc, float* Lsen_Gsc) {
  if (!d || !vel || !depz || !tRc || !pvRc || !Lsen_Gsc)
    return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (d->nx < 1 || d->ny < 1 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d", d->nx, d->ny, d->nz,
                   d->kmax);
  if (d->vp_rho_mode != 0)
    return set_err(DAZI_DK_EINVAL, "dazi_dk_ti derives vp and rho from vs (vp_rho_mode 0)");
  return host_stage(d, vel, nullptr, nullptr, depz, tRc, pvRc, nullptr, nullptr, nullptr,
                    Lsen_Gsc);
}

int32_t dazi_dk_iso(const dazi_dk_desc* d, const float* vel, const float* vp, const float* rho,
                    const float* depz, const double* tRc, double* pvRc, double* sen_vsRc,
                    double* sen_vpRc, double* sen_rhoRc) {
  if (!d || !vel || !depz || !tRc || !pvRc) return set_err(DAZI_DK_EINVAL, "NULL required argument");
  if (d->nx < 1 || d->ny < 1 || d->nz < 2 || d->kmax < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes nx=%d ny=%d nz=%d kmax=%d", d->nx, d->ny, d->nz,
                   d->kmax);
  if (d->vp_rho_mode == 1 && (!vp || !rho))
    return set_err(DAZI_DK_EINVAL, "vp_rho_mode=1 needs vp and rho");
  const bool kern = sen_vsRc || sen_vpRc || sen_rhoRc;
  if (kern && !(sen_vsRc && sen_vpRc && sen_rhoRc))
    return set_err(DAZI_DK_EINVAL, "sen_vs, sen_vp, sen_rho must be all set or all NULL");
  return host_stage(d, vel, vp, rho, depz, tRc, pvRc, sen_vsRc, sen_vpRc, sen_rhoRc, nullptr);
}

int32_t dazi_dk_layered_dev(const dazi_dk_layered* b, const dazi_dk_tune* tune, void* stream) {
  bool m32 = b && b->thk && b->vp && b->vs && b->rho;
  bool m64 = b && b->thk64 && b->vp64 && b->vs64 && b->rho64;
  if (!b || !b->nl || !b->loff || !b->nk || !b->toff || !(m32 || m64) || !b->periods || !b->c ||
      !b->status)
    return set_err(DAZI_DK_EINVAL, "NULL required argument in dazi_dk_layered");
  if (b->ncol < 0 || b->max_layers < 1 || b->nsolves < 0)
    return set_err(DAZI_DK_EINVAL, "bad sizes ncol=%d max_layers=%d", b->ncol, b->max_layers);
  if (b->ncol == 0) return 0;
  dazi_dk_tune T;
  if (int rc = check_tune(tune, &T)) return rc;
  bool kern = b->dc_dvs || b->dc_dvp || b->dc_drho;
  if (kern && !b->solve_col) return set_err(DAZI_DK_EINVAL, "kernels need solve_col");
  hipStream_t st = (hipStream_t)stream;
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  int64_t nlay_total = b->nlayers_total;
  if (nlay_total <= 0) {
    // total layers: loff[ncol-1] + nl[ncol-1] read back from the device
    int32_t nl_last = 0;
    int64_t lo_last = 0;
    HIPCHK(hipMemcpyAsync(&nl_last, b->nl + (b->ncol - 1), 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&lo_last, b->loff + (b->ncol - 1), 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    nlay_total = lo_last + nl_last;
  }
  void *layers, *btp, *cc, *betmx, *bad, *croot, *work;
  if (int rc = pool_get(P, S_LAYERS, (size_t)nlay_total * 32, &layers)) return rc;
  if (int rc = pool_get(P, S_BTP, (size_t)nlay_total * 16, &btp)) return rc;
  double* btr = (double*)btp + nlay_total;
  if (int rc = pool_get(P, S_CC, (size_t)b->ncol * 8, &cc)) return rc;
  if (int rc = pool_get(P, S_BETMX, (size_t)b->ncol * 4, &betmx)) return rc;
  if (int rc = pool_get(P, S_BAD, (size_t)b->ncol * 4, &bad)) return rc;
  if (b->c_unrounded) {
    croot = b->c_unrounded;
  } else if (int rc = pool_get(P, S_CROOT, (size_t)(b->nsolves > 0 ? b->nsolves : 1) * 8,
                               &croot)) {
    return rc;
  }
  if (int rc = pool_get(P, S_WORK, 64, &work)) return rc;
  // scheduling order: columns (and, for the kernel pass, solves) by layer
  // count, largest first
  void *order, *sorder = nullptr;
  if (int rc = pool_get(P, S_ORDER, (2 * (size_t)b->ncol + dk::order_work_ints()) * 4, &order))
    return rc;
  if (kern && b->nsolves > 0) {
    if (int rc = pool_get(P, S_SORDER, (size_t)b->nsolves * 4, &sorder)) return rc;
  }
  P->gen++;
  P->have_events = g_timing;
  P->last_L = 0;
  P->last_ti_threads = 0;
  rec(P, 0, st);
  HIPCHK(dk::launch_prep_layers(b->ncol, b->nl, b->loff, b->thk, b->vp, b->vs, b->rho,
                                b->thk64, b->vp64, b->vs64, b->rho64, (double4*)layers,
                                (double*)btp, btr, (double*)cc, (float*)betmx, (int*)bad, st));
  // S_ORDER: the column order, each column's first solve position, the counters
  HIPCHK(dk::launch_order(b->ncol, b->max_layers, b->nl, b->nk, b->toff, b->solve_col, b->nsolves,
                          (int*)order, (int*)sorder, (int*)order + b->ncol,
                          (int*)order + 2 * (size_t)b->ncol, st));
  rec(P, 1, st);
  dk::BatchDev B = {};
  B.ncol = b->ncol;
  B.uniform = 0;
  B.L = b->nl;
  B.loff = b->loff;
  B.K = b->nk;
  B.toff = b->toff;
  B.layers = (const double4*)layers;
  B.btp = (const double*)btp;
  B.btr = btr;
  B.periods = b->periods;
  B.cc = (const double*)cc;
  B.betmx = (const float*)betmx;
  B.bad = (const int*)bad;
  B.croot = (double*)croot;
  B.pv = b->c;
  B.status = b->status;
  B.work = (int*)work;
  B.order = (const int*)order;
  B.tol = T.refine_tol;
  if (int rc = run_forward(P, B, T, b->max_layers, st)) return rc;
  rec(P, 2, st);
  if (kern) {
    dk::KernDev O = {};
    O.node_mode = 0;
    O.nsolves = b->nsolves;
    O.solve_col = b->solve_col;
    O.solve_order = (const int*)sorder;
    O.dvs = b->dc_dvs;
    O.dvp = b->dc_dvp;
    O.drho = b->dc_drho;
    int nthr = kernel_threads(P, b->nsolves);
    int L = b->max_layers;
    void* tape;
    if (int rc = pool_get(P, S_TAPE, dk::kern_tape_bytes(L, nthr), &tape))
      return rc;
    HIPCHK(dk::launch_kernels(B, O, (double*)tape, dk::kern_checkpoints(L), nthr, st));
  }
  rec(P, 3, st);
  rec(P, 4, st);  // no TI phase
  return 0;
}

}  // extern "C"

// ---- the allocation-free root search
//
// dazi_dk_layered_noalloc_dev keeps its work memory (flattened layers, their
// flattening factors, the per-column start velocity, betmx and prep flags, the
// unrounded roots when the caller passes none, and the persistent work
// counter) in this file-scope device array instead of the pool.  Every device
// the library runs on holds its own copy until the process ends.  On the host
// the array's name is only a stand-in: host code takes the device address that
// hipGetSymbolAddress returns for the current device (NoAllocDev::ptr) and
// passes that to the kernels, never the name itself.  The array keeps the last
// call's contents; the work counter, the one value that must start at zero, is
// cleared by a kernel at every call.  Calls that overlap on one device race on
// the array.
static __device__ __align__(256) unsigned char scratch[DAZI_DK_NOALLOC_SCRATCH_BYTES];

namespace {

struct NoAllocDev {
  void* ptr = nullptr;  // device address of `scratch` on this device
  int ncu = 0;
};
NoAllocDev g_noalloc[16];

// Address of the array and CU count of the current device, looked up on the
// first call on that device (no allocation: the array is part of the loaded
// code object).
int noalloc_device(NoAllocDev** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err(DAZI_DK_ENODEV, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= 16) return set_err(DAZI_DK_ENODEV, "device index %d unsupported", dev);
  NoAllocDev* N = &g_noalloc[dev];
  if (!N->ptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(DAZI_DK_EHIP, "hipGetSymbolAddress(scratch) failed: %s",
                     hipGetErrorString(e));
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    N->ncu = ncu > 0 ? ncu : 1;
    N->ptr = p;
  }
  *out = N;
  return 0;
}

inline size_t align256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" {

int32_t dazi_dk_layered_noalloc_dev(const dazi_dk_layered* b, const dazi_dk_tune* tune,
                                    void* stream) {
  bool m32 = b && b->thk && b->vp && b->vs && b->rho;
  bool m64 = b && b->thk64 && b->vp64 && b->vs64 && b->rho64;
  if (!b || !b->nl || !b->loff || !b->nk || !b->toff || !(m32 || m64) || !b->periods || !b->c ||
      !b->status)
    return set_err(DAZI_DK_EINVAL, "NULL required argument in dazi_dk_layered");
  if (b->dc_dvs || b->dc_dvp || b->dc_drho)
    return set_err(DAZI_DK_EINVAL,
                   "dazi_dk_layered_noalloc_dev computes phase velocities only (kernel outputs "
                   "must be NULL)");
  if (b->ncol < 0 || b->max_layers < 1 || b->nsolves < 0)
    return set_err(DAZI_DK_EINVAL, "bad sizes ncol=%d max_layers=%d", b->ncol, b->max_layers);
  if (b->nlayers_total <= 0)
    return set_err(DAZI_DK_EINVAL, "dazi_dk_layered_noalloc_dev needs nlayers_total");
  if (b->ncol == 0) return 0;
  dazi_dk_tune T;
  if (int rc = check_tune(tune, &T)) return rc;
  // the array's layout: work counter, per-column values, per-layer records,
  // then the unrounded roots unless the caller gives c_unrounded
  const size_t ncol = (size_t)b->ncol, nlay = (size_t)b->nlayers_total;
  const size_t o_cc = 256, o_betmx = o_cc + align256(ncol * 8),
               o_bad = o_betmx + align256(ncol * 4), o_lay = o_bad + align256(ncol * 4),
               o_btp = o_lay + align256(nlay * 32), o_croot = o_btp + align256(nlay * 16);
  const size_t need =
      o_croot + (b->c_unrounded ? 0 : align256((size_t)(b->nsolves > 0 ? b->nsolves : 1) * 8));
  if (need > (size_t)DAZI_DK_NOALLOC_SCRATCH_BYTES)
    return set_err(DAZI_DK_ELIMIT,
                   "batch needs %zu bytes of work memory, the allocation-free entry point has "
                   "%zu (%zu columns, %zu layers, %lld solves)",
                   need, (size_t)DAZI_DK_NOALLOC_SCRATCH_BYTES, ncol, nlay,
                   (long long)b->nsolves);
  NoAllocDev* N;
  if (int rc = noalloc_device(&N)) return rc;
  {
    dazi_dk_tune R = T;  // the LDS staging of the deepest model fits (device limit)
    if (int rc = resolve_shape(&R, b->max_layers, b->ncol, N->ncu)) return rc;
  }
  unsigned char* base = (unsigned char*)N->ptr;
  hipStream_t st = (hipStream_t)stream;
  double4* layers = (double4*)(base + o_lay);
  double* btp = (double*)(base + o_btp);
  double* btr = btp + nlay;
  double* cc = (double*)(base + o_cc);
  float* betmx = (float*)(base + o_betmx);
  int* bad = (int*)(base + o_bad);
  double* croot = b->c_unrounded ? b->c_unrounded : (double*)(base + o_croot);
  HIPCHK(dk::launch_prep_layers(b->ncol, b->nl, b->loff, b->thk, b->vp, b->vs, b->rho, b->thk64,
                                b->vp64, b->vs64, b->rho64, layers, btp, btr, cc, betmx, bad, st));
  dk::BatchDev B = {};
  B.ncol = b->ncol;
  B.uniform = 0;
  B.L = b->nl;
  B.loff = b->loff;
  B.K = b->nk;
  B.toff = b->toff;
  B.layers = layers;
  B.btp = btp;
  B.btr = btr;
  B.periods = b->periods;
  B.cc = cc;
  B.betmx = betmx;
  B.bad = bad;
  B.croot = croot;
  B.pv = b->c;
  B.status = b->status;
  B.work = (int*)base;
  B.tol = T.refine_tol;
  int nlaunch = 0;
  return forward_launches(B, T, b->max_layers, N->ncu, st, &nlaunch);
}

int32_t dazi_dk_layered_vjp_dev(const dazi_dk_layered* b, const double* grad_c, double* grad_vs,
                                double* grad_vp, double* grad_rho, double* grad_thk,
                                void* stream) {
  bool m32 = b && b->thk && b->vp && b->vs && b->rho;
  bool m64 = b && b->thk64 && b->vp64 && b->vs64 && b->rho64;
  if (!b || !b->nl || !b->loff || !b->nk || !b->toff || !(m32 || m64) || !b->periods ||
      !b->c_unrounded || !grad_c || !grad_vs || !grad_vp || !grad_rho)
    return set_err(DAZI_DK_EINVAL, "NULL required argument (the VJP needs c_unrounded)");
  if (b->ncol < 0 || b->max_layers < 1)
    return set_err(DAZI_DK_EINVAL, "bad sizes ncol=%d max_layers=%d", b->ncol, b->max_layers);
  if (b->ncol == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  Pool* P;
  if (int rc = get_pool(&P)) return rc;
  int64_t nlay_total = b->nlayers_total;
  if (nlay_total <= 0) {
    int32_t nl_last = 0;
    int64_t lo_last = 0;
    HIPCHK(hipMemcpyAsync(&nl_last, b->nl + (b->ncol - 1), 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&lo_last, b->loff + (b->ncol - 1), 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    nlay_total = lo_last + nl_last;
  }
  void *layers, *btp, *cc, *betmx, *bad, *tape;
  if (int rc = pool_get(P, S_LAYERS, (size_t)nlay_total * 32, &layers)) return rc;
  if (int rc = pool_get(P, S_BTP, (size_t)nlay_total * 16, &btp)) return rc;
  double* btr = (double*)btp + nlay_total;
  if (int rc = pool_get(P, S_CC, (size_t)b->ncol * 8, &cc)) return rc;
  if (int rc = pool_get(P, S_BETMX, (size_t)b->ncol * 4, &betmx)) return rc;
  if (int rc = pool_get(P, S_BAD, (size_t)b->ncol * 4, &bad)) return rc;
  int nblocks = P->ncu * 8;
  if (nblocks > b->ncol) nblocks = b->ncol;
  int L = b->max_layers;
  if (int rc = pool_get(P, S_TAPE, (size_t)6 * (L > 1 ? L - 1 : 1) * nblocks * 64 * 8, &tape))
    return rc;
  P->gen++;
  P->have_events = g_timing;
  rec(P, 0, st);
  HIPCHK(dk::launch_prep_layers(b->ncol, b->nl, b->loff, b->thk, b->vp, b->vs, b->rho,
                                b->thk64, b->vp64, b->vs64, b->rho64, (double4*)layers,
                                (double*)btp, btr, (double*)cc, (float*)betmx, (int*)bad, st));
  rec(P, 1, st);
  dk::BatchDev B = {};
  B.ncol = b->ncol;
  B.uniform = 0;
  B.L = b->nl;
  B.loff = b->loff;
  B.K = b->nk;
  B.toff = b->toff;
  B.layers = (const double4*)layers;
  B.btp = (const double*)btp;
  B.btr = btr;
  B.periods = b->periods;
  B.croot = b->c_unrounded;
  HIPCHK(dk::launch_vjp(B, grad_c, grad_vs, grad_vp, grad_rho, grad_thk, m64 ? nullptr : b->thk,
                        m64 ? b->thk64 : nullptr, (double*)tape, nblocks, st));
  rec(P, 2, st);
  rec(P, 3, st);
  rec(P, 4, st);
  return 0;
}

}  // extern "C"
