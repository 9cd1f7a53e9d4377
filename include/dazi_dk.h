/*
 * dazi_dk.h — C ABI of the MI355X depth-kernel stage.
 *
 * The stage computes, for every grid column and period, the Rayleigh
 * fundamental-mode phase velocity and its depth sensitivity kernels.  It is a
 * drop-in for the reference's two Fortran subroutines
 *
 *   depthkernel(nx,ny,nz,vel,pvRc,sen_vsRc,sen_vpRc,sen_rhoRc,
 *               iwave,igr,kmaxRc,tRc,depz,minthk)
 *       code/src_inv_iso_joint/CalSurfG.f90:1-15   (ISO kernels, FD backend :86-107)
 *   depthkernelTI(nx,ny,nz,vel,pvRc,iwave,igr,kmaxRc,tRc,depz,minthk,Lsen_Gsc)
 *       code/src_inv_iso_joint/depthkernelTI.f90:2-34 (TI dc/dL splice :152-162)
 *
 * All arrays use the reference's Fortran column-major layout: the column index
 * post = ii + (jj-1)*nx is the fastest dimension of every output.  Entry points
 * taking host pointers are synchronous; entry points with the _dev suffix take
 * device pointers and enqueue on the given HIP stream.  Calls from several host
 * threads at once are not supported (the reference is not reentrant either).
 *
 * Device work buffers belong to a per-device pool that only grows; buffers a
 * growth replaces are kept until the process ends, so a HIP graph captured
 * around a device entry point stays valid after later, larger calls, and
 * after calls with another node geometry (each geometry's sublayer counts and
 * layer records live in buffers of their own).
 *
 * Return codes: 0 ok; otherwise one of DAZI_DK_E* (the Fortran shim turns a
 * non-zero code into `stop <code>`), and dazi_dk_last_error() describes it.
 * A period whose root is not found is not an error: as in the reference
 * (surfdisp96.f:307-348) that period and all later ones of the column are 0.
 */
#ifndef DAZI_DK_H
#define DAZI_DK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DAZI_DK_ABI_VERSION 1

#define DAZI_DK_OK 0
#define DAZI_DK_EINVAL 9301      /* invalid argument or unsupported option */
#define DAZI_DK_EHIP 9302        /* HIP runtime error */
#define DAZI_DK_ELIMIT 9303      /* size beyond a documented limit */
#define DAZI_DK_ENODEV 9304      /* no usable GPU */

/* Stage description: the reference's scalar arguments plus the para.in tile
 * controls (pdsurf_tile_columns / ti_tile_columns). */
typedef struct dazi_dk_desc {
  int32_t nx, ny, nz;      /* columns = nx*ny; nz depth nodes */
  int32_t kmax;            /* number of periods (kmaxRc) */
  int32_t iwave;           /* 2 = Rayleigh: the only value the reference stage uses */
  int32_t igr;             /* 0 = phase velocity: the only value the reference stage uses */
  int32_t tile_columns;    /* columns per launch in plain mode (tune.mode = 1), unless an
                              explicit tune argument sets its own; 0 = 64.  The default
                              persistent mode runs the whole batch in one launch */
  int32_t vp_rho_mode;     /* 0: vp, rho from Brocher(vs) as depthkernel does;
                              1: explicit vp and rho node arrays */
  float minthk0;           /* para.in "number of sublayers" (minthk) */
  int32_t flags;           /* 0, or DAZI_DK_DEPZ_UNCHANGED */
} dazi_dk_desc;

/* desc->flags bit for the device entry points: the caller guarantees that the
 * values behind depz are those of the previous call with the same depz
 * pointer, so the library may skip reading them back (a host synchronisation)
 * to check its cached node geometry. */
#define DAZI_DK_DEPZ_UNCHANGED 1

/* desc->flags bit for the TI splice: periods without a phase velocity get NaN
 * in every interval of Lsen_Gsc, as the reference's depthkernelTI produces
 * there (it calls tregn96 with c = 0, depthkernelTI.f90:139-146); without it
 * they get 0.  Either way the G builder drops them (abs(NaN) > ftol is false,
 * CalSurfGAniso_Joint.f90:782-789). */
#define DAZI_DK_TI_FAILED_NAN 2

/* Launch configuration of the root search. Zero fields take defaults. */
typedef struct dazi_dk_tune {
  int32_t lanes_per_scan;  /* trial velocities per scan step: 8, 16, 32 or 64; 0 (default):
                              the most of 16, 32 and 64 that keep the launch within one
                              wave per SIMD (64 up to 1,024 columns on 256 CUs, 16 from
                              2,049 on) */
  int32_t cols_per_wg;     /* columns per workgroup (lanes_per_scan*cols_per_wg <= 256);
                              0 (default): 256-thread workgroups, fewer columns where the
                              deepest model's LDS staging (64 B per layer and column) would
                              exceed the device's LDS per workgroup */
  int32_t mode;            /* 0: batched persistent launch (default);
                              1: plain, one launch per tile of tile_columns columns */
  int32_t tile_columns;    /* plain-mode tile (default 64 = pdsurf_tile_columns) */
  double refine_tol;       /* root bracket width before the final secant step, km/s
                              (default 2e-6) */
} dazi_dk_tune;

/* Ragged batch of layered models, all pointers on the device.  Column j has
 * nl[j] layers at loff[j] of thk/vp/vs/rho (half-space last, its thickness is
 * ignored) and nk[j] ascending periods at toff[j] of periods.  Outputs are per
 * solve s = toff[j]+k: c (float32-rounded phase velocity, 0 where not found)
 * and, when non-NULL, refined-layer kernels dc/dVs, dc/dVp, dc/drho of the
 * unflattened model at kern[s + nsolves*m] (m < nl[j]; rows beyond are
 * untouched).  status[j] = 0, or the 1-based first failed period, or -1 for a
 * model the stage does not handle (fluid layer). */
typedef struct dazi_dk_layered {
  int32_t ncol;
  int32_t max_layers;      /* max over nl[] */
  int64_t nsolves;         /* sum of nk[] */
  const int32_t* nl;
  const int64_t* loff;
  const int32_t* nk;
  const int64_t* toff;
  const int32_t* solve_col; /* column of each solve (needed only for kernels) */
  const float* thk;
  const float* vp;
  const float* vs;
  const float* rho;
  const double* periods;
  double* c;
  int32_t* status;
  double* dc_dvs;
  double* dc_dvp;
  double* dc_drho;
  double* c_unrounded;     /* optional (may be NULL): the float64 root before the
                              float32 rounding of c, 0 where not found */
  const double* thk64;     /* optional float64 model, used instead of thk/vp/vs/rho
                              when all four are set (flattened in float64; for the */
  const double* vp64;      /* differentiable path: the reference's own models are */
  const double* vs64;      /* float32).  Kernels and VJPs are then with respect to */
  const double* rho64;     /* these float64 values. */
  int64_t nlayers_total;   /* sum of nl[] when > 0; 0 makes the call read it back
                              from the device (a host synchronisation) */
} dazi_dk_layered;

int32_t dazi_dk_abi_version(void);
const char* dazi_dk_last_error(void);

/* Default launch configuration used by every entry point without a tune
 * argument (the Fortran/host path). */
int32_t dazi_dk_set_tune(const dazi_dk_tune* tune);
int32_t dazi_dk_get_tune(dazi_dk_tune* tune);

/* Replaces depthkernel (CalSurfG.f90:1-15); host pointers, synchronous.
 *   vel(nx,ny,nz) f32; vp/rho (nx,ny,nz) f32 only when desc->vp_rho_mode = 1;
 *   depz(nz) f32; tRc(kmax) f64 ascending;
 *   pvRc(nx*ny,kmax) f64; sen_*(nx*ny,kmax,nz) f64 node kernels (may all be
 *   NULL for phase velocities only). */
int32_t dazi_dk_iso(const dazi_dk_desc* desc, const float* vel, const float* vp,
                    const float* rho, const float* depz, const double* tRc, double* pvRc,
                    double* sen_vsRc, double* sen_vpRc, double* sen_rhoRc);

/* Same stage on device pointers, enqueued on `stream` (a hipStream_t, NULL =
 * default stream).  status(nx*ny) i32 may be NULL.  tune may be NULL.
 * The refined node geometry (sublayer counts, shared layer records) is derived
 * from the values of depz and kept on the device.  A later call with the same
 * (depz pointer, nz, minthk) reads the nz depths back (a host
 * synchronisation) and reuses the geometry when they are unchanged; with
 * desc->flags = DAZI_DK_DEPZ_UNCHANGED, or on a stream being captured, it only
 * enqueues kernels, so after one such call an identical call can be captured
 * in a HIP graph.  A captured graph keeps the geometry it was captured with:
 * when a replay finds depz holding other values, every column gets status -2
 * and zero outputs.  Status per column: 0, the 1-based first period without
 * a root, -1 for a model outside the stage's class (fluid or non-positive
 * layer), -2 as just described. */
int32_t dazi_dk_iso_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                        const float* vp, const float* rho, const float* depz, const double* tRc,
                        double* pvRc, double* sen_vsRc, double* sen_vpRc, double* sen_rhoRc,
                        int32_t* status, void* stream);

/* Makes the next device call on the current device read depz back and look
 * its values up among the node geometries met (no reference counterpart: the
 * reference recomputes the geometry on every call).  Not needed for
 * correctness: without DAZI_DK_DEPZ_UNCHANGED every eager call compares. */
int32_t dazi_dk_reset_geometry(void);

/* Replaces depthkernelTI (depthkernelTI.f90:2-34); host pointers, synchronous.
 * vp and rho come from Brocher(vs) as in the reference (desc->vp_rho_mode 0).
 *   pvRc(nx*ny,kmax) f64 phase velocities (the reference recomputes them too);
 *   Lsen_Gsc(nx*ny,kmax,nz-1) f32 = sum over the sublayers of interval j of
 *   A dc/dA + L dc/dL (tregn96 kernels at the float32 phase velocity). */
int32_t dazi_dk_ti(const dazi_dk_desc* desc, const float* vel, const float* depz,
                   const double* tRc, double* pvRc, float* Lsen_Gsc);

/* TI stage on device pointers (vp/rho explicit when desc->vp_rho_mode = 1). */
int32_t dazi_dk_ti_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                       const float* vp, const float* rho, const float* depz, const double* tRc,
                       double* pvRc, float* Lsen_Gsc, int32_t* status, void* stream);

/* Both subroutines' outputs from one root search (the G builder calls
 * depthkernelTI and depthkernel back to back on the same model,
 * CalSurfGAniso_Joint.f90:478,486).  sen_* may be NULL (all three) and
 * Lsen_Gsc may be NULL. */
int32_t dazi_dk_stage_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune, const float* vel,
                          const float* vp, const float* rho, const float* depz, const double* tRc,
                          double* pvRc, double* sen_vsRc, double* sen_vpRc, double* sen_rhoRc,
                          float* Lsen_Gsc, int32_t* status, void* stream);

/* Refined layers per column (half-space included) of the node geometry
 * (depz(nz), minthk), computed on the host as CalSurfG.f90:2506-2520 does: the
 * first dimension of dazi_dk_ti_kernels_dev's per-layer outputs. */
int32_t dazi_dk_node_layers(int32_t nz, const float* depz, float minthk, int32_t* L);

/* The TI stage of dazi_dk_ti_dev with, in addition, the per-refined-layer
 * kernels that depthkernelTI splices into Lsen_Gsc (depthkernelTI.f90:152-162):
 *   dc_dA(m, k, col) = dc/dalpha_h / (2 rho vp) - F/(A - 2L)^2 dc/deta,
 *   dc_dL(m, k, col) = dc/dbeta_v / (2 rho vs) + 2F/(A - 2L)^2 dc/deta,
 * f32, (L, kmax, nx*ny) with the column fastest, for all L refined layers
 * (dazi_dk_node_layers; the half-space last, which the splice leaves out),
 * from tregn96's kernels at the float32 phase velocity.  Lsen_Gsc may be
 * NULL.  No reference counterpart as an interface: the reference forms these
 * values inside depthkernelTI and keeps only their splice. */
int32_t dazi_dk_ti_kernels_dev(const dazi_dk_desc* desc, const dazi_dk_tune* tune,
                               const float* vel, const float* vp, const float* rho,
                               const float* depz, const double* tRc, double* pvRc,
                               float* Lsen_Gsc, float* dc_dA, float* dc_dL, int32_t* status,
                               void* stream);

/* Vector-Jacobian product over the TI kernels: for every refined layer m and
 * column, grad_A(m, col) = sum_k grad_c(k, col) dc_dA(m, k, col) and grad_L
 * likewise (f64, (L, ncol)); periods whose kernels are not finite contribute
 * nothing.  Deterministic.  grad_A or grad_L may be NULL.  No reference
 * counterpart (the reference has no gradient code). */
int32_t dazi_dk_ti_vjp_dev(int32_t ncol, int32_t L, int32_t kmax, const float* dc_dA,
                           const float* dc_dL, const double* grad_c, double* grad_A,
                           double* grad_L, void* stream);

/* Ragged layered batch (the batched persistent solver's native input). */
int32_t dazi_dk_layered_dev(const dazi_dk_layered* batch, const dazi_dk_tune* tune,
                            void* stream);

/* Bytes of the device work memory of dazi_dk_layered_noalloc_dev (a static
 * array that every device the library runs on holds until the process ends). */
#define DAZI_DK_NOALLOC_SCRATCH_BYTES (32u << 20)

/* The ragged batch's root search without any device allocation: the same
 * phase velocities c, c_unrounded (optional) and status as
 * dazi_dk_layered_dev, bit for bit, with its work memory in the library's
 * static array of DAZI_DK_NOALLOC_SCRATCH_BYTES instead of the growing pool.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The kernel outputs dc_dvs/dc_dvp/dc_drho must be
 * NULL and nlayers_total must be set.  Limit: 256 + 48 B per layer + 16 B per
 * column (+ 8 B per solve when c_unrounded is NULL), each part rounded up to
 * 256 B, within DAZI_DK_NOALLOC_SCRATCH_BYTES; a larger batch returns
 * DAZI_DK_ELIMIT before any device work.  All calls on one device share the
 * array, so calls that overlap on a device (other streams, other host
 * threads) race. */
int32_t dazi_dk_layered_noalloc_dev(const dazi_dk_layered* batch, const dazi_dk_tune* tune,
                                    void* stream);

/* Vector-Jacobian product of the layered batch's phase velocities: for every
 * layer m of every column, grad_x[loff+m] = sum over the column's periods of
 * grad_c[s] * dc_s/dx_m, x in {vs, vp, rho, thk} of the unflattened model, by
 * implicit differentiation of the secular function at the root
 * (dc/dx = -(dF/dx)/(dF/dc)); the thickness gradient includes the earth
 * flattening of every deeper layer (surfdisp96.f:480-547) and is 0 for the
 * half-space.  grad_thk may be NULL.  batch->c_unrounded must hold the roots of
 * a dazi_dk_layered_dev call on the same batch; periods without a root
 * contribute nothing.  Deterministic (no atomics).  No reference counterpart
 * (the reference differentiates by finite differences, DepthKernel1D.f90:146-183). */
int32_t dazi_dk_layered_vjp_dev(const dazi_dk_layered* batch, const double* grad_c,
                                double* grad_vs, double* grad_vp, double* grad_rho,
                                double* grad_thk, void* stream);

/* Effective dc/dVs of the G builder's Vs row (CalSurfGAniso_Joint.f90:750-765):
 * with vp and rho tied to vs by Brocher,
 *   eff(z,k,col) = sen_vs + sen_vp*coe_a + sen_rho*coe_rho, z < nz-1 (the
 *   reference's rows 1:nz-1), coefficients in float32 from vel(col, z);
 * and, when grad_vs is set, its contraction with grad_c(col, k):
 *   grad_vs(col, z) = sum_k grad_c * eff for every node z (the VJP of c with
 *   respect to node Vs under the Brocher relation).
 * Device pointers in the stage's layouts: vel (nx*ny, nz) f32; sen_*
 * (nx*ny, kmax, nz) f64; grad_c (nx*ny, kmax); eff (nx*ny, kmax, nz-1) f64 or
 * NULL; grad_vs (nx*ny, nz) f64 or NULL. */
int32_t dazi_dk_effective_dev(const dazi_dk_desc* desc, const float* vel, const double* sen_vs,
                              const double* sen_vp, const double* sen_rho, const double* grad_c,
                              double* eff, double* grad_vs, void* stream);

/* Per-cell sums of a population gradient (the Voronoi realizations' cell
 * gradient; no counterpart in the reference, whose Voronoi projection is
 * outside this library): for the chunk of columns c in [c0, c0+n) of nreal
 * realizations of ncol_real columns each (c = r*ncol_real + j),
 *   acc[(r*nl + l)*ncells + cell[(r*nl + l)*ncol_real + j]] += gl[l*ld + c - c0]
 * for every layer l < nl.  Device pointers: gl (nl, ld) f64, cell (nreal, nl,
 * ncol_real) i64, acc (nreal, nl, ncells) f64.  Each sum runs in column order
 * (deterministic, no atomics); cell ids outside [0, ncells) are skipped.
 * Calls on one stream accumulate in order. */
int32_t dazi_dk_cell_sum_dev(const double* gl, int64_t ld, int32_t nl, int64_t c0, int64_t n,
                             int32_t ncol_real, int32_t nreal, const int64_t* cell,
                             int32_t ncells, double* acc, void* stream);

/* G-row assembly: the sparse rows (rw, iw, col) of the joint inversion's G
 * matrix that the reference builds ray by ray from the path weights and the
 * stage outputs (CalSurfGAniso_Joint.f90:750-797, inside the source/period/
 * record loops :511-799; replaces that loop body's row fill and compression).
 * For ray r (data row iw = row_base + r, 1-based) of period index ray_k[r]
 * (0-based, the reference's knumi - 1) and each of its cells with
 * |fdm| >= ftol, the row holds over z = 0..nz-2
 *   Vs block: real((sen_vp*coe_a + sen_rho*coe_rho + sen_vs)(post,k,z) * fdm)
 *   Gc block: Lsen_Gsc(post,k,z) * fdmc,   Gs block: Lsen_Gsc(post,k,z) * fdms
 * at model column col = block*nparpi + z*nvx*nvz + cell + 1 (1-based),
 * nparpi = nvx*nvz*(nz-1), post = (cell/nvx + 1)*(nvx+2) + cell%nvx + 1; the
 * entries with |value| > ftol are emitted in increasing (iw, col), exactly the
 * reference's order.  A ray's cells are given as cell = (jj-1)*nvx + (kk-1)
 * (the interior cell of fdm(jj,kk)), strictly increasing, in
 * cell[cell_off[r] .. cell_off[r+1]) with their fdm, fdmc, fdms.  The stage
 * arrays keep the stage's layouts over the full nx*ny = (nvx+2)*(nvz+2) grid:
 * vel (nx*ny, nz) f32, sen_* (nx*ny, kmax, nz) f64, lsen (nx*ny, kmax, nz-1)
 * f32.  ftol is the reference's 1e-4. */
typedef struct dazi_dk_grow_desc {
  int32_t nvx, nvz;        /* interior grid: nx-2, ny-2 */
  int32_t nz;              /* depth nodes of the stage arrays */
  int32_t kmax;            /* periods of the stage arrays */
  int64_t nray;            /* rays (data rows) in this call */
  int32_t row_base;        /* iw of ray 0 (the reference's count1 of that ray) */
  float ftol;              /* 1e-4 in the reference */
} dazi_dk_grow_desc;

/* Device pointers; enqueues on stream.  Writes the row count of each ray to
 * row_nnz (device, nray; NULL = internal) and the total to *nar (device
 * int64), and the first min(total, cap) entries to rw, iw, col (device).  A
 * caller that needs all entries sizes cap from *nar (an upper bound is
 * 3*(nz-1)*cell_off[nray]).  Rays with an invalid period or cell index get no
 * entries for it (dazi_dk_grows rejects them). */
int32_t dazi_dk_grows_dev(const dazi_dk_grow_desc* desc, const int32_t* ray_k,
                          const int64_t* cell_off, const int32_t* cell, const float* fdm,
                          const float* fdmc, const float* fdms, const float* vel,
                          const double* sen_vs, const double* sen_vp, const double* sen_rho,
                          const float* lsen, int64_t cap, float* rw, int32_t* iw, int32_t* col,
                          int64_t* row_nnz, int64_t* nar, void* stream);

/* Host pointers; synchronous.  Validates the ray description (period and cell
 * ranges, increasing cells), uploads, assembles, and returns the entries with
 * their count in *nar; DAZI_DK_ELIMIT (with *nar set) when they exceed cap. */
int32_t dazi_dk_grows(const dazi_dk_grow_desc* desc, const int32_t* ray_k,
                      const int64_t* cell_off, const int32_t* cell, const float* fdm,
                      const float* fdmc, const float* fdms, const float* vel,
                      const double* sen_vs, const double* sen_vp, const double* sen_rho,
                      const float* lsen, int64_t cap, float* rw, int32_t* iw, int32_t* col,
                      int64_t* nar);

/* Period azimuthal-anisotropy map: the forward 2-psi anisotropy of every
 * interior cell and period from Lsen_Gsc and the Gc, Gs layer models
 * (replaces the computation of FwdAzimuthalAniMap,
 * code/src_inv_iso_joint/FwdAzimuthalAniMap.f90:44-89; its unit-42 text
 * output stays with the caller):
 *   cos_rc = sum_k Lsen_Gsc(post, t, k) * gcf(ii, jj, k),  sin_rc likewise with gsf,
 *   amp = 0.5*sqrt(cos^2 + sin^2),  amp_rela = amp / real(tRcV(cell, t)),
 *   angle = 0.5*(atan2(sin, cos)/pi*180, +360 when negative)  [degrees],
 * post = jj*nx + ii (1-based interior ii, jj; 0-based column), in float32 as
 * the reference.  Layouts: lsen (nx*ny, kmax, nz-1) f32, gcf and gsf
 * (nx-2, ny-2, nz-1) f32, tRcV ((nx-2)*(ny-2), kmax) f64, outputs
 * ((nx-2)*(ny-2), kmax) f32, all Fortran order.  angle agrees with the
 * reference to 2 units in the last place (its libm atan2f is not correctly
 * rounded); the other outputs are exact. */
typedef struct dazi_dk_azim_desc {
  int32_t nx, ny, nz, kmax;
} dazi_dk_azim_desc;

/* Device pointers; enqueues on stream.  amp, amp_rela, angle may be NULL
 * (tRcV then too, when amp_rela is NULL). */
int32_t dazi_dk_azimuth_map_dev(const dazi_dk_azim_desc* desc, const float* lsen,
                                const float* gcf, const float* gsf, const double* tRcV,
                                float* cos_rc, float* sin_rc, float* amp, float* amp_rela,
                                float* angle, void* stream);

/* Host pointers, all outputs required; synchronous. */
int32_t dazi_dk_azimuth_map(const dazi_dk_azim_desc* desc, const float* lsen, const float* gcf,
                            const float* gsf, const double* tRcV, float* cos_rc, float* sin_rc,
                            float* amp, float* amp_rela, float* angle);

/* First-arrival traveltimes: the consumer of the phase-velocity maps pvRc.
 * Replaces the traveltime part of the joint G builder's source loop
 * (code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:511-699 and the srtimes
 * calls of :720-725): for every (period, source) job, gridder
 * (CalSurfG.f90:1596-1689) dices the period's map onto the propagation grid,
 * bsplrefine (:1698-1770) and travel (:399-598, fouds1/fouds2 :608-870, the
 * binary tree :879-1032) march the first arrivals on the refined source grid
 * and then on the coarse grid, and srtimes (:1772-1906) gives the time at each
 * of the job's receivers.  Results are bit-identical to the reference's
 * single-precision routines.
 *
 * Grid constants are those CalSurfGAniso_Joint.f90:375-409 sets (gdx = gdz =
 * 5, sgdl = sgs = 8, asgr = 1, fom = 1, snb = 0.5, earth = 6371; nvx = nx-2,
 * nvz = ny-2; goxd, gozd, dvxd, dvzd as given to the G builder).  The
 * propagation grid has nnx = (nvx-1)*gdx+1 by nnz = (nvz-1)*gdz+1 nodes. */
typedef struct dazi_dk_travel_desc {
  int32_t nvx, nvz;   /* inversion grid without its border */
  int32_t gdx, gdz;   /* propagation-grid dicing */
  int32_t sgdl, sgs;  /* source-grid dicing level, its half-width in coarse nodes */
  int32_t asgr, fom;  /* source-grid refinement (0/1), mixed-order scheme (0/1) */
  float goxd, gozd;   /* latitude and longitude of the grid origin, degrees */
  float dvxd, dvzd;   /* inversion-grid spacing, degrees */
  float snb;          /* narrow-band capacity as a fraction of the grid's nodes */
  float earth;        /* Earth radius, km */
  int32_t kmax;       /* periods in pvRc */
  int32_t slots;      /* concurrent jobs on the device (0: automatic) */
} dazi_dk_travel_desc;

typedef struct dazi_dk_travel_plan dazi_dk_travel_plan;

/* Host arrays in: njob jobs, job q at period job_k[q] (0-based column of
 * pvRc) with its source at colatitude scx[q] and longitude scz[q] (radians,
 * as scxf/sczf of the G builder) and receivers rec_off[q] ..
 * rec_off[q+1]-1 of rcx/rcz (radians).  Validates (the reference stops on a
 * source or receiver outside the model), computes the sin tables the
 * marching reads, uploads the geometry and reserves the per-slot work memory
 * on the current device.  The plan belongs to that device. */
int32_t dazi_dk_travel_plan_create(const dazi_dk_travel_desc* desc, int64_t njob,
                                   const int32_t* job_k, const float* scx, const float* scz,
                                   const int64_t* rec_off, const float* rcx, const float* rcz,
                                   dazi_dk_travel_plan** plan);

/* Device pointers; enqueues on stream.  pvRc (nx*ny, kmax) f64 as the stage
 * writes it; times (rec_off[njob]) f32 in job then receiver order (the
 * reference's dsurf order); ttn (njob, nnz*nnx) f32 coarse fields in the
 * reference's ttn(nnz, nnx) layout, or NULL; status (njob) i32: 0, or 1 when
 * the source falls outside its refined grid (the reference stops), 2 when a
 * job's narrow band outgrows snb*nodes (the reference would write past its
 * tree), 3 when the refined pass stopped before its band reached a coarse node
 * (a source within about a coarse node of the south or east edge: the
 * reference's exit test at CalSurfG.f90:503-523 compares the coarse vnr/vnb
 * with the refined node count, and its coarse field stays unset), 4 when a
 * receiver in the source cell meets a source on the last grid node (the
 * reference reads past the grid); such jobs' times are NaN.
 * Calls on one plan must not overlap. */
int32_t dazi_dk_travel_run_dev(dazi_dk_travel_plan* plan, const double* pvRc, float* times,
                               float* ttn, int32_t* status, void* stream);

int32_t dazi_dk_travel_plan_destroy(dazi_dk_travel_plan* plan);

/* Host pointers; synchronous: plan, upload pvRc, run, copy back, destroy.
 * ttn may be NULL. */
int32_t dazi_dk_travel(const dazi_dk_travel_desc* desc, int64_t njob, const int32_t* job_k,
                       const float* scx, const float* scz, const int64_t* rec_off,
                       const float* rcx, const float* rcz, const double* pvRc, float* times,
                       float* ttn, int32_t* status);

/* Bytes of device memory the library's pool holds on the current device:
 * current slots, buffers retired by growth, and the buffers of every node
 * geometry met (one set per distinct depz values and minthk). */
int32_t dazi_dk_pool_bytes(int64_t* out);

/* Event timings (ms) of the last call's phases on its stream:
 * [0] model prep, [1] root search, [2] isotropic kernel pass, [3] TI kernels
 * (their model prep runs on a second stream beside the root search), [4]
 * number of root-search launches, [5] refined layers per column of the last
 * node-model call (0 for a ragged batch), [6] device threads of its TI
 * launch (0 without TI).  Waits for those events.  Returns the number of
 * values written (at most n, at most 7). */
int32_t dazi_dk_last_timings(double* out, int32_t n);

/* Enables (1) or disables (0) recording of the timings above. */
int32_t dazi_dk_enable_timing(int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* DAZI_DK_H */
