// Internal device-side descriptors shared by the kernels and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dk {

// Column status codes besides 0 and the 1-based first failed period: a model
// outside the stage's class, and (graph replays only) a depz whose values
// differ from those the captured node geometry was built from.
enum { DK_STATUS_BADMODEL = -1, DK_STATUS_GEOMETRY = -2 };
// Prep flag per column (BatchDev::bad): 0 ok, 1 bad model, 2 changed geometry.
enum { DK_PREP_BADMODEL = 1, DK_PREP_GEOMETRY = 2 };

// A batch of columns with refined, flattened layered models.
//   uniform = 1: every column has Lu layers (at col*Lu) and the Ku shared
//                periods; solve index s = k*ncol + col (column fastest, the
//                reference's pvRc(nx*ny,kmax) layout).
//   uniform = 0: column col has L[col] layers at loff[col] and K[col] periods
//                at toff[col]; solve index s = toff[col] + k.
struct BatchDev {
  int ncol;
  int uniform;
  int Lu, Ku;
  const int* L;
  const int64_t* loff;
  const int* K;
  const int64_t* toff;
  const double4* layers;  // (d, a, b, rho) flattened
  const double* btp;      // flattening velocity factor per layer
  const double* btr;      // btp^-2.275 per layer (rho's flattening factor)
  const double* periods;
  const double* cc;      // start velocity (= cm) per column
  const float* betmx;    // max flattened shear velocity per column
  const int* bad;        // prep status per column (0 ok)
  double* croot;         // refined root per solve (0 where not found)
  double* pv;            // float32-rounded phase velocity per solve
  int* status;           // 0, or 1-based first failed period, or -1 bad model
  int* work;             // persistent work counter
  const int* order;      // ragged batches: the columns in scheduling order
                         // (position w of a launch's range solves column
                         // order[w]); NULL: index order
  double tol;            // refinement bracket width (km/s)
};

// Kernel-pass outputs.
//   node_mode = 1: node kernels in the reference layout sen(ncol, kmax, nz)
//                  (needs uniform batches and nsub per node interval);
//   node_mode = 0: refined-layer kernels at s + nsolves*m.
struct KernDev {
  int node_mode;
  int nz;
  const int* nsub;
  int64_t nsolves;
  const int* solve_col;  // ragged batches: column of each solve
  const int* solve_order;  // ragged batches: solves in scheduling order (NULL:
                           // index order)
  double* dvs;
  double* dvp;
  double* drho;
};

// One refined layer of the TI model (stored by k_prep_ti in word planes: word
// w of (layer m, column col) at [(m * 15 + w) * ncol + col]): unflattened
// values used by the splice
// (d, rvp, rvs, rrho, TA, TL, TF) and flattened ones used by the eigen-solution
// (df, A = C, L = N, F, rho) plus the sphericity factor vtp; float32 as in the
// reference.  k holds the layer's period-independent FP64 constants of the
// eigen-solution and energy integrals, in ti::Layer's order: 1/C, 1/L,
// eta = F/(A-2L), F/C, sqrt(A/rho), sqrt(C/rho), sqrt(L/rho), F/eta.
struct TiLayerDev {
  float d, rvp, rvs, rrho, TA, TL, TF;
  float df, A, L, F, rho, vtp, pad;
  double k[8];
};

// Doubles of TI scratch per layer and thread: real solves, complex-pair solves.
constexpr int TI_SCRATCH = 9;
constexpr int TI_SCRATCH_CX = 14;

// Kernel pass: layers per checkpoint segment.  The tangent sweep keeps the
// state entering every KERN_SEG-th layer (5 doubles and an int exponent per
// checkpoint and thread); the adjoint sweep recomputes a segment's states into LDS
// (KERN_SEG x (5 doubles + 1 int) per thread: 78,848 B for 256 threads, so
// two 256-thread blocks share a CU's 160 KiB).
constexpr int KERN_SEG = 7;
constexpr int KERN_BLOCK = 256;
constexpr int KERN_LDS = KERN_SEG * KERN_BLOCK * (5 * 8 + 4);
// checkpoints per thread for models of at most L layers (half-space included),
// and the checkpoint buffer's bytes for nthr threads
inline int kern_checkpoints(int L) { return L > 1 ? (L - 1 + KERN_SEG - 1) / KERN_SEG : 1; }
inline size_t kern_tape_bytes(int L, int64_t nthr) {
  return (size_t)kern_checkpoints(L) * (size_t)nthr * (5 * 8 + 4);
}

// k_prep_ti (node model and the shared records of launch_prep_geom) also
// clears queue[0]; launch_ti runs k_ti, which appends the solves
// with a complex-conjugate eigenvalue pair to queue[1..], then k_ti_cx over
// them (queue holds 1 + ncol * K ints).
hipError_t launch_prep_ti(int ncol, const float* vel, const float* vp, const float* rho, int L,
                          const void* geo, TiLayerDev* out, int* queue, hipStream_t st);
// nan_failed: periods without a phase velocity get NaN in every interval
// (the reference's tregn96 at c = 0) instead of 0.
// lsen may be NULL; dcdA/dcdL (both or neither) receive the refined-layer
// kernels dc/dA, dc/dL at ((m*K + k)*ncol + col).
hipError_t launch_ti(int ncol, int K, int L, int nz, const TiLayerDev* layers,
                     const double* periods, const double* pv, const int* nsub, float* lsen,
                     float* dcdA, float* dcdL, double* scratch, int nthr, int* queue,
                     int nan_failed, hipStream_t st);
hipError_t launch_ti_vjp(int ncol, int L, int K, const float* dcdA, const float* dcdL,
                         const double* gc, double* gA, double* gL, hipStream_t st);

// launch_prep_geom writes the L flattening and node-mapping records that
// every column of a node geometry (depz, minthk0) shares to geo
// (PREP_GEOM_BYTES each); launch_prep_nodes reads them for the columns' layers.
constexpr size_t PREP_GEOM_BYTES = 80;
// launch_prep_geom also copies depz to geo_depz; launch_prep_nodes compares
// depz with that copy and flags every column DK_PREP_GEOMETRY on a mismatch.
hipError_t launch_prep_geom(int nz, const float* depz, float minthk0, int L, void* geo,
                            float* geo_depz, hipStream_t st);
hipError_t launch_prep_nodes(int ncol, int nz, const float* vel, const float* vp, const float* rho,
                             const float* depz, const float* geo_depz, int L, const void* geo,
                             double4* layers, double* btp, double* btr, double* cc, float* betmx,
                             int* bad, hipStream_t st);
hipError_t launch_prep_layers(int ncol, const int* Lc, const int64_t* loff, const float* thk,
                              const float* vp, const float* vs, const float* rho,
                              const double* thk64, const double* vp64, const double* vs64,
                              const double* rho64, double4* layers, double* btp, double* btr,
                              double* cc, float* betmx, int* bad, hipStream_t st);
hipError_t launch_forward(const BatchDev& B, int S, int col_begin, int col_end, int cols_per_wg,
                          int persistent, int grid, int lds_stride, hipStream_t st);
int forward_max_blocks(int S, int cols_per_wg, int lds_stride);
hipError_t launch_kernels(const BatchDev& B, const KernDev& O, double* tape, int nq,
                          int grid_threads, hipStream_t st);
hipError_t launch_zero_int(int* p, int n, hipStream_t st);
// Scheduling order of a ragged batch: its columns, and (solve_order non-NULL)
// its solves, sorted by the column's layer count and then its period count,
// largest first, so that the groups of a wavefront and the threads of the
// kernel pass run columns of nearly equal depth and length, and a launch ends
// on its shortest columns.  A counting sort over ORDER_BINS keys: the layer
// count (those from ORDER_BINS - 1 up share the last value) and, where the
// keys leave room for it below max_layers, up to 4 bits of the period count;
// work holds order_work_ints() ints.  The order within a key is unspecified
// (no result depends on it).  The solve order needs solve_col and spos, ncol
// ints of work (each column's first position in it).
constexpr int ORDER_BINS = 4096;
inline int order_work_ints() { return 4 * ORDER_BINS; }
hipError_t launch_order(int ncol, int max_layers, const int* Lc, const int* Kc,
                        const int64_t* toff, const int* solve_col, int64_t nsolves, int* order,
                        int* solve_order, int* spos, int* work, hipStream_t st);
// G-row assembly (CalSurfGAniso_Joint.f90:750-797), device pointers; see
// dazi_dk_grows_dev.  row_nnz, row_off: per ray; nar: one int64.
struct GrowDev {
  int nvx, nvz, nz, K;
  int64_t nray;
  int row_base;
  float ftol;
  const int* ray_k;
  const int64_t* cell_off;
  const int* cell;
  const float* fdm;
  const float* fdmc;
  const float* fdms;
  const float* vel;
  const double* svs;
  const double* svp;
  const double* srho;
  const float* lsen;
  double* effT;  // work: effective dc/dVs, (K, ncol, nz-1) depth fastest
  float* lsenT;  // work: Lsen_Gsc, same layout
  int64_t* tsum;  // work: grow_scan_tiles(nray) tile sums
  int64_t cap;
  float* rw;
  int* iw;
  int* col;
  int64_t* row_nnz;
  int64_t* row_off;
  int64_t* nar;
};
hipError_t launch_grows(const GrowDev& g, hipStream_t st);
long grow_scan_tiles(long nray);

// Period azimuthal-anisotropy map (FwdAzimuthalAniMap.f90:44-89); see
// dazi_dk_azimuth_map_dev.
hipError_t launch_azimuth_map(int nvx, int nvz, int nz, int K, const float* lsen, const float* gcf,
                              const float* gsf, const double* tRcV, float* cos_rc, float* sin_rc,
                              float* amp, float* amp_rela, float* angle, hipStream_t st);

// First-arrival traveltimes of (period, source) jobs on the propagation grid
// (CalSurfGAniso_Joint.f90:511-725); see dazi_dk_travel_run_dev.
struct TravelDev {
  // grid (CalSurfGAniso_Joint.f90:375-409)
  int nvx, nvz, gdx, gdz, sgdl, sgs, asgr, fom;
  int nnx, nnz, K;
  long ncol;
  float gox, goz, dnx, dnz, drnx, drnz, earth, snb;
  float dpl;           // srtimes' near-source distance threshold
  int maxbt_c;         // tree capacity nint(snb*nnx*nnz)
  int nrx_max, nr_max;  // refined-grid bounds (nodes)
  long nrp_max;         // refined-grid node records with their halo
  const double* pv;    // (ncol, K) phase velocities, column fastest
  float* veln;         // work: (K, nnz*nnx) diced velocities
  const float* risti_c;  // nnx: earth*sinf(gox + (ix-1)*dnx)
  const float* risti_r;  // (nnx, nrx_max): the same on the refined grid starting at column vnl
  // jobs and receivers
  long njob;
  const int* job_k;
  const float* scx;
  const float* scz;
  const int64_t* rec_off;
  const float* rcx;
  const float* rcz;
  const float* sin_rcx;
  // outputs
  float* times;
  float* ttn;  // optional (njob, nnz*nnx)
  int* status;
  // per-slot work: node records with a 3-node halo on every side
  int2* node_c;  // (nnx+6)*(nnz+6) per slot
  int2* node_r;  // nrp_max per slot
  float* vel_r;
  int2* spill;
  long spill_cap;
  int slots;
};
hipError_t launch_travel(const TravelDev& p, hipStream_t st);

hipError_t launch_effective(int ncol, int nz, int K, const float* vel, const double* svs,
                            const double* svp, const double* srho, const double* gc, double* eff,
                            double* grad, hipStream_t st);
// Per-cell sums of a population gradient; see dazi_dk_cell_sum_dev.
hipError_t launch_cell_sum(const double* gl, int64_t ld, int nl, int64_t c0, int64_t n, int ncr,
                           const int64_t* cell, int ncells, double* acc, hipStream_t st);
hipError_t launch_vjp(const BatchDev& B, const double* grad_c, double* gvs, double* gvp,
                      double* grho, double* gthk, const float* thk32, const double* thk64,
                      double* tape, int nblocks, hipStream_t st);

}  // namespace dk
