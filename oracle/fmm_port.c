/* Test infrastructure only: a CPU restatement of the reference's first-arrival
 * traveltime stage, the consumer of the phase-velocity maps pvRc.  Only tests/,
 * the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * One job = one (period, source) pair of the joint G builder
 * (code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:511-725):
 *   gridder     CalSurfG.f90:1596-1689   cubic B-spline dicing of the period's
 *                                        (nvz+2) x (nvx+2) control values onto
 *                                        the nnz x nnx propagation grid
 *   bsplrefine  CalSurfG.f90:1698-1770   the same on the refined source grid
 *   travel      CalSurfG.f90:399-598     fast marching with a binary-tree
 *                                        narrow band (addtree :879-925,
 *                                        downtree :927-1003, updtree :1005-1032)
 *   fouds1/2    CalSurfG.f90:608-696, 698-870  first-order / mixed-order
 *                                        upwind updates (spherical grid)
 *   srtimes     CalSurfG.f90:1772-1906   receiver times (bilinear, or the
 *                                        straight-ray estimate near the source)
 * with the grid set-up of CalSurfGAniso_Joint.f90:375-432 and the source-grid
 * refinement sequence of :526-699.
 *
 * Arithmetic is single precision throughout (globalp's i10 is
 * SELECTED_REAL_KIND(6)), evaluated in the reference's operation order with
 * contraction off; integer powers are repeated products; sin is libm sinf.
 * Arrays are column-major as in the reference: node (iz, ix), 1-based, at
 * (iz-1) + (ix-1)*nz.
 *
 * Parity: pinned bit-for-bit against the reference's own routines compiled
 * from CalSurfG.f90 (oracle/_ref, driven by oracle/fmm_ref_driver.f90).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#pragma STDC FP_CONTRACT OFF

typedef struct {
    int nvx, nvz, gdx, gdz, sgdl, sgs, asgr, fom;
    float goxd, gozd, dvxd, dvzd, snb, earth;
} orc_fmm_desc;

/* One grid (coarse or refined) and its fast-marching state. */
typedef struct {
    int nx, nz;               /* nnx, nnz */
    float gox, goz, dnx, dnz;
    const float *vel;         /* veln(nz, nx) */
    float *t;                 /* ttn */
    int *st;                  /* nsts: -1 far, 0 alive, >0 tree position */
    int16_t *bx, *bz;         /* btg(:)%px, %pz (1-based) */
    int ntr, maxbt, overflow;
    float earth;
    int fom;
    /* refined-grid exit rule (urg = 1) */
    int vnl, vnr, vnt, vnb;
} grid_t;

static const float PI_F = 3.1415926535898f; /* globalp's pi, single precision */

#define IX(g, iz, ix) ((size_t)((iz) - 1) + (size_t)((ix) - 1) * (size_t)(g)->nz)

static float p2(float x) { return x * x; }
static float p3(float x) { return x * x * x; }

/* ---- binary tree (narrow band), CalSurfG.f90:879-1032 ---- */

static float key(const grid_t *g, int p) { return g->t[IX(g, g->bz[p], g->bx[p])]; }

static void swap_entries(grid_t *g, int a, int b) {
    int16_t x = g->bx[a], z = g->bz[a];
    g->bx[a] = g->bx[b];
    g->bz[a] = g->bz[b];
    g->bx[b] = x;
    g->bz[b] = z;
}

static void sift_up(grid_t *g, int iz, int ix, int tpc) {
    int tpp = tpc / 2;
    float tv = g->t[IX(g, iz, ix)];
    while (tpp > 0) {
        if (tv < key(g, tpp)) {
            g->st[IX(g, iz, ix)] = tpp;
            g->st[IX(g, g->bz[tpp], g->bx[tpp])] = tpc;
            swap_entries(g, tpc, tpp);
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
}

static void addtree(grid_t *g, int iz, int ix) {
    if (g->ntr + 1 > g->maxbt) { /* the reference writes past btg here */
        g->overflow = 1;
        return;
    }
    g->ntr++;
    g->st[IX(g, iz, ix)] = g->ntr;
    g->bx[g->ntr] = (int16_t)ix;
    g->bz[g->ntr] = (int16_t)iz;
    sift_up(g, iz, ix, g->ntr);
}

static void updtree(grid_t *g, int iz, int ix) { sift_up(g, iz, ix, g->st[IX(g, iz, ix)]); }

static void set_pos(grid_t *g, int p, int v) { g->st[IX(g, g->bz[p], g->bx[p])] = v; }

static void downtree(grid_t *g) {
    if (g->ntr == 1) {
        g->ntr = 0;
        return;
    }
    set_pos(g, g->ntr, 1);
    g->bx[1] = g->bx[g->ntr];
    g->bz[1] = g->bz[g->ntr];
    g->ntr--;
    int tpp = 1, tpc = 2;
    while (tpc < g->ntr) {
        if (key(g, tpc) > key(g, tpc + 1)) tpc++;
        if (key(g, tpc) < key(g, tpp)) {
            set_pos(g, tpp, tpc);
            set_pos(g, tpc, tpp);
            swap_entries(g, tpc, tpp);
            tpp = tpc;
            tpc = 2 * tpp;
        } else {
            tpc = g->ntr + 1;
        }
    }
    if (tpc == g->ntr) {
        if (key(g, tpc) < key(g, tpp)) {
            set_pos(g, tpp, tpc);
            set_pos(g, tpc, tpp);
            swap_entries(g, tpc, tpp);
        }
    }
}

/* ---- upwind updates ---- */

/* rd1 = b^2 - 4ac clamped at 0; (-b + sqrt(rd1)) / (2a) */
static float quad_root(float a, float b, float c) {
    float rd1 = p2(b) - 4.0f * a * c;
    if (rd1 < 0.0f) rd1 = 0.0f;
    return (-b + sqrtf(rd1)) / (2.0f * a);
}

/* fouds1, CalSurfG.f90:608-696 */
static void fouds1(grid_t *g, int iz, int ix) {
    int tsw1 = 0;
    float travm = 0.0f;
    float slown = 1.0f / g->vel[IX(g, iz, ix)];
    float ri = g->earth;
    float risti = ri * sinf(g->gox + (float)(ix - 1) * g->dnx);
    for (int j = ix - 1; j <= ix + 1; j += 2) {
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (j < 1 || j > g->nx || k < 1 || k > g->nz) continue;
            int sol = 0;
            float a = 0, b = 0, c = 0, tref = 0;
            if (g->st[IX(g, iz, j)] == 0) {
                sol = 1;
                if (g->st[IX(g, k, ix)] == 0) {
                    float u = ri * g->dnx, v = risti * g->dnz;
                    float em = g->t[IX(g, k, ix)] - g->t[IX(g, iz, j)];
                    a = p2(u) + p2(v);
                    b = -2.0f * p2(u) * em;
                    c = p2(u) * (p2(em) - p2(v) * p2(slown));
                    tref = g->t[IX(g, iz, j)];
                } else {
                    a = 1.0f;
                    b = 0.0f;
                    c = -p2(slown) * p2(ri) * p2(g->dnx);
                    tref = g->t[IX(g, iz, j)];
                }
            } else if (g->st[IX(g, k, ix)] == 0) {
                sol = 1;
                a = 1.0f;
                b = 0.0f;
                c = -p2(slown * risti * g->dnz);
                tref = g->t[IX(g, k, ix)];
            }
            if (sol) {
                float trav = tref + quad_root(a, b, c);
                if (tsw1) {
                    travm = fminf(trav, travm);
                } else {
                    travm = trav;
                    tsw1 = 1;
                }
            }
        }
    }
    g->t[IX(g, iz, ix)] = travm;
}

/* fouds2, CalSurfG.f90:698-870 */
static void fouds2(grid_t *g, int iz, int ix) {
    int tsw1 = 0;
    float travm = 0.0f;
    float slown = 1.0f / g->vel[IX(g, iz, ix)];
    float ri = g->earth;
    float risti = ri * sinf(g->gox + (float)(ix - 1) * g->dnx);
    const float *T = g->t;
    for (int j = ix - 1; j <= ix + 1; j += 2) {
        if (j < 1 || j > g->nx) continue;
        int j2 = (j == ix - 1) ? j - 1 : j + 1;
        int swj = -1;
        if (j2 >= 1 && j2 <= g->nx && g->st[IX(g, iz, j2)] == 0) swj = 0;
        if (g->st[IX(g, iz, j)] == 0 && swj == 0) {
            swj = (T[IX(g, iz, j)] > T[IX(g, iz, j2)]) ? 0 : -1;
        } else {
            swj = -1;
        }
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (k < 1 || k > g->nz) continue;
            int k2 = (k == iz - 1) ? k - 1 : k + 1;
            int swk = -1;
            if (k2 >= 1 && k2 <= g->nz && g->st[IX(g, k2, ix)] == 0) swk = 0;
            if (g->st[IX(g, k, ix)] == 0 && swk == 0) {
                swk = (T[IX(g, k, ix)] > T[IX(g, k2, ix)]) ? 0 : -1;
            } else {
                swk = -1;
            }
            int sol = 0;
            float a = 0, b = 0, c = 0, tref = 0, tdiv = 1.0f, u, v, em;
            int ajz = g->st[IX(g, iz, j)] == 0, akx = g->st[IX(g, k, ix)] == 0;
            if (swj == 0) {
                sol = 1;
                if (swk == 0) {
                    u = 2.0f * ri * g->dnx;
                    v = 2.0f * risti * g->dnz;
                    em = 4.0f * T[IX(g, iz, j)] - T[IX(g, iz, j2)] - 4.0f * T[IX(g, k, ix)];
                    em = em + T[IX(g, k2, ix)];
                    a = p2(v) + p2(u);
                    b = 2.0f * em * p2(u);
                    c = p2(u) * (p2(em) - p2(slown) * p2(v));
                    tref = 4.0f * T[IX(g, iz, j)] - T[IX(g, iz, j2)];
                    tdiv = 3.0f;
                } else if (akx) {
                    u = risti * g->dnz;
                    v = 2.0f * ri * g->dnx;
                    em = 3.0f * T[IX(g, k, ix)] - 4.0f * T[IX(g, iz, j)] + T[IX(g, iz, j2)];
                    a = p2(v) + 9.0f * p2(u);
                    b = 6.0f * em * p2(u);
                    c = p2(u) * (p2(em) - p2(slown) * p2(v));
                    tref = T[IX(g, k, ix)];
                    tdiv = 1.0f;
                } else {
                    u = 2.0f * ri * g->dnx;
                    a = 1.0f;
                    b = 0.0f;
                    c = -p2(u) * p2(slown);
                    tref = 4.0f * T[IX(g, iz, j)] - T[IX(g, iz, j2)];
                    tdiv = 3.0f;
                }
            } else if (ajz) {
                sol = 1;
                if (swk == 0) {
                    u = ri * g->dnx;
                    v = 2.0f * risti * g->dnz;
                    em = 3.0f * T[IX(g, iz, j)] - 4.0f * T[IX(g, k, ix)] + T[IX(g, k2, ix)];
                    a = p2(v) + 9.0f * p2(u);
                    b = 6.0f * em * p2(u);
                    c = p2(u) * (p2(em) - p2(v) * p2(slown));
                    tref = T[IX(g, iz, j)];
                    tdiv = 1.0f;
                } else if (akx) {
                    u = ri * g->dnx;
                    v = risti * g->dnz;
                    em = T[IX(g, k, ix)] - T[IX(g, iz, j)];
                    a = p2(u) + p2(v);
                    b = -2.0f * p2(u) * em;
                    c = p2(u) * (p2(em) - p2(v) * p2(slown));
                    tref = T[IX(g, iz, j)];
                    tdiv = 1.0f;
                } else {
                    a = 1.0f;
                    b = 0.0f;
                    c = -p2(slown) * p2(ri) * p2(g->dnx);
                    tref = T[IX(g, iz, j)];
                    tdiv = 1.0f;
                }
            } else {
                if (swk == 0) {
                    sol = 1;
                    u = 2.0f * risti * g->dnz;
                    a = 1.0f;
                    b = 0.0f;
                    c = -p2(u) * p2(slown);
                    tref = 4.0f * T[IX(g, k, ix)] - T[IX(g, k2, ix)];
                    tdiv = 3.0f;
                } else if (akx) {
                    sol = 1;
                    a = 1.0f;
                    b = 0.0f;
                    c = -p2(slown) * p2(risti) * p2(g->dnz);
                    tref = T[IX(g, k, ix)];
                    tdiv = 1.0f;
                }
            }
            if (sol) {
                float trav = (tref + quad_root(a, b, c)) / tdiv;
                if (tsw1) {
                    travm = fminf(trav, travm);
                } else {
                    travm = trav;
                    tsw1 = 1;
                }
            }
        }
    }
    g->t[IX(g, iz, ix)] = travm;
}

static void update(grid_t *g, int iz, int ix) {
    if (g->fom == 0) fouds1(g, iz, ix);
    else fouds2(g, iz, ix);
}

/* bilinear, CalSurfG.f90:2466-2488: nv[i][j] = node (isz-1+j, isx-1+i) */
static float bilinear(float nv[2][2], float dsx, float dsz, float dnx, float dnz) {
    float biv = 0.0f;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            float produ = (1.0f - fabsf(((float)(i - 1) * dnx - dsx) / dnx)) *
                          (1.0f - fabsf(((float)(j - 1) * dnz - dsz) / dnz));
            biv = biv + nv[i - 1][j - 1] * produ;
        }
    return biv;
}

/* travel, CalSurfG.f90:399-598.  urg: 0 plain, 1 refined source grid (exit when
 * the band reaches an edge of the refined grid that is not a model edge), 2
 * start from the band already marked in st. */
static int travel(grid_t *g, float scx, float scz, int urg) {
    int isx = (int)((scx - g->gox) / g->dnx) + 1;
    int isz = (int)((scz - g->goz) / g->dnz) + 1;
    if (isx < 1 || isx > g->nx || isz < 1 || isz > g->nz) return 1;
    if (isx == g->nx) isx--;
    if (isz == g->nz) isz--;
    size_t n = (size_t)g->nx * g->nz;
    if (urg != 2)
        for (size_t i = 0; i < n; i++) g->st[i] = -1;
    g->ntr = 0;
    if (urg == 2) {
        for (int i = 1; i <= g->nx; i++)
            for (int j = 1; j <= g->nz; j++)
                if (g->st[IX(g, j, i)] > 0) addtree(g, j, i);
    } else {
        float vss[2][2];
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = g->vel[IX(g, isz - 1 + j, isx - 1 + i)];
        float dsx = (scx - g->gox) - (float)(isx - 1) * g->dnx;
        float dsz = (scz - g->goz) - (float)(isz - 1) * g->dnz;
        float vsrc = bilinear(vss, dsx, dsz, g->dnx, g->dnz);
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) {
                float ds = sqrtf(p2(dsx - (float)(i - 1) * g->dnx) + p2(dsz - (float)(j - 1) * g->dnz));
                g->t[IX(g, isz - 1 + j, isx - 1 + i)] = 2.0f * ds / (vss[i - 1][j - 1] + vsrc);
                addtree(g, isz - 1 + j, isx - 1 + i);
            }
    }
    while (g->ntr > 0 && !g->overflow) {
        int ix = g->bx[1], iz = g->bz[1];
        if (urg == 1) {
            int swrg = 0;
            if (ix == 1 && g->vnl != 1) swrg = 1;
            if (ix == g->nx && g->vnr != g->nx) swrg = 1;
            if (iz == 1 && g->vnt != 1) swrg = 1;
            if (iz == g->nz && g->vnb != g->nz) swrg = 1;
            if (swrg) {
                g->st[IX(g, iz, ix)] = 0;
                break;
            }
        }
        g->st[IX(g, iz, ix)] = 0;
        downtree(g);
        for (int i = ix - 1; i <= ix + 1; i += 2) {
            if (i < 1 || i > g->nx) continue;
            int s = g->st[IX(g, iz, i)];
            if (s == -1) {
                update(g, iz, i);
                addtree(g, iz, i);
            } else if (s > 0) {
                update(g, iz, i);
                updtree(g, iz, i);
            }
        }
        for (int i = iz - 1; i <= iz + 1; i += 2) {
            if (i < 1 || i > g->nz) continue;
            int s = g->st[IX(g, i, ix)];
            if (s == -1) {
                update(g, i, ix);
                addtree(g, i, ix);
            } else if (s > 0) {
                update(g, i, ix);
                updtree(g, i, ix);
            }
        }
    }
    return g->overflow ? 2 : 0;
}

/* B-spline basis at parameter (i-1)/n, CalSurfG.f90:1644-1651 */
static void basis(int i, int n, float w[4]) {
    float u = (float)n;
    u = (float)(i - 1) / u;
    w[0] = p3(1.0f - u) / 6.0f;
    w[1] = (4.0f - 6.0f * p2(u) + 3.0f * p3(u)) / 6.0f;
    w[2] = (1.0f + 3.0f * u + 3.0f * p2(u) - 3.0f * p3(u)) / 6.0f;
    w[3] = p3(u) / 6.0f;
}

/* velv(i, j), i = 0..nvz+1, j = 0..nvx+1, from pv (column i*(nvx+2)+j) */
#define VV(i, j) velv[(size_t)(i) * (size_t)(d->nvx + 2) + (size_t)(j)]

/* gridder, CalSurfG.f90:1596-1689 */
void orc_fmm_gridder(const orc_fmm_desc *d, const double *pv, float *veln) {
    int nvx = d->nvx, nvz = d->nvz, nnz = (nvz - 1) * d->gdz + 1;
    size_t nc = (size_t)(nvx + 2) * (nvz + 2);
    float *velv = malloc(nc * sizeof(float));
    for (size_t c = 0; c < nc; c++) velv[c] = (float)pv[c];
    for (int i = 1; i <= nvz - 1; i++) {
        int conz = (i == nvz - 1) ? d->gdz + 1 : d->gdz;
        for (int j = 1; j <= nvx - 1; j++) {
            int conx = (j == nvx - 1) ? d->gdx + 1 : d->gdx;
            for (int l = 1; l <= conz; l++) {
                int stz = d->gdz * (i - 1) + l;
                float vi[4];
                basis(l, d->gdz, vi);
                for (int m = 1; m <= conx; m++) {
                    int stx = d->gdx * (j - 1) + m;
                    float ui[4];
                    basis(m, d->gdx, ui);
                    float sumi = 0.0f;
                    for (int i1 = 1; i1 <= 4; i1++) {
                        float sumj = 0.0f;
                        for (int j1 = 1; j1 <= 4; j1++) sumj = sumj + ui[j1 - 1] * VV(i - 2 + i1, j - 2 + j1);
                        sumi = sumi + vi[i1 - 1] * sumj;
                    }
                    veln[(size_t)(stz - 1) + (size_t)(stx - 1) * nnz] = sumi;
                }
            }
        }
    }
    free(velv);
}

/* bsplrefine, CalSurfG.f90:1698-1770: the refined grid over coarse nodes
 * [vnl, vnr] x [vnt, vnb]; out is nrz x nrx. */
static void bsplrefine(const orc_fmm_desc *d, const float *velv, int vnl, int vnr, int vnt, int vnb,
                       int nrx, int nrz, float *out) {
    int nvx = d->nvx, nvz = d->nvz;
    int nrxr = d->gdx * d->sgdl, nrzr = d->gdz * d->sgdl;
    int origx = (vnl - 1) * d->sgdl + 1, origz = (vnt - 1) * d->sgdl + 1;
    for (int i = 1; i <= nvz - 1; i++) {
        int conz = (i == nvz - 1) ? nrzr + 1 : nrzr;
        for (int j = 1; j <= nvx - 1; j++) {
            int conx = (j == nvx - 1) ? nrxr + 1 : nrxr;
            for (int k = 1; k <= conz; k++) {
                int st1 = d->gdz * (i - 1) + (k - 1) / d->sgdl + 1;
                if (st1 < vnt || st1 > vnb) continue;
                st1 = nrzr * (i - 1) + k;
                float vi[4];
                basis(k, nrzr, vi);
                for (int l = 1; l <= conx; l++) {
                    int st2 = d->gdx * (j - 1) + (l - 1) / d->sgdl + 1;
                    if (st2 < vnl || st2 > vnr) continue;
                    st2 = nrxr * (j - 1) + l;
                    float ui[4], sum[4];
                    basis(l, nrxr, ui);
                    for (int i1 = 1; i1 <= 4; i1++) {
                        sum[i1 - 1] = 0.0f;
                        for (int j1 = 1; j1 <= 4; j1++)
                            sum[i1 - 1] = sum[i1 - 1] + ui[j1 - 1] * VV(i - 2 + i1, j - 2 + j1);
                        sum[i1 - 1] = vi[i1 - 1] * sum[i1 - 1];
                    }
                    int idm1 = st1 - origz + 1, idm2 = st2 - origx + 1;
                    if (idm1 < 1 || idm1 > nrz || idm2 < 1 || idm2 > nrx) continue;
                    out[(size_t)(idm1 - 1) + (size_t)(idm2 - 1) * nrz] = sum[0] + sum[1] + sum[2] + sum[3];
                }
            }
        }
    }
}

/* Coarse grid constants, CalSurfGAniso_Joint.f90:396-409 */
static void coarse_grid(const orc_fmm_desc *d, float *gox, float *goz, float *dnx, float *dnz, float *dvx,
                        float *dvz) {
    *dvx = d->dvxd * PI_F / 180.0f;
    *dvz = d->dvzd * PI_F / 180.0f;
    *gox = (90.0f - d->goxd) * PI_F / 180.0f;
    *goz = d->gozd * PI_F / 180.0f;
    *dnx = *dvx / (float)d->gdx;
    *dnz = *dvz / (float)d->gdz;
}

/* srtimes, CalSurfG.f90:1772-1906, on the coarse grid c.  Returns 1 when the
 * receiver lies outside the model (the reference stops). */
static int srtimes(const grid_t *c, float scx, float scz, float rcx, float rcz, float *out) {
    int irx = (int)((rcx - c->gox) / c->dnx) + 1;
    int irz = (int)((rcz - c->goz) / c->dnz) + 1;
    if (irx < 1 || irx > c->nx || irz < 1 || irz > c->nz) return 1;
    if (irx == c->nx) irx--;
    if (irz == c->nz) irz--;
    int isx = (int)((scx - c->gox) / c->dnx) + 1;
    int isz = (int)((scz - c->goz) / c->dnz) + 1;
    float dpl = c->dnx * c->earth;
    float rd1 = c->dnz * c->earth * sinf(c->gox);
    if (rd1 < dpl) dpl = rd1;
    rd1 = c->dnz * c->earth * sinf(c->gox + (float)(c->nx - 1) * c->dnx);
    if (rd1 < dpl) dpl = rd1;
    float sred = p2((scx - rcx) * c->earth);
    sred = sred + p2((scz - rcz) * c->earth * sinf(rcx));
    sred = sqrtf(sred);
    int sw = 0;
    if (sred < dpl) sw = 1;
    if (isx == irx && isz == irz) sw = 1;
    float trr;
    if (sw) {
        /* the reference reads the source cell at (isx, isz) without the
         * last-node adjustment; a source on the last node column or row would
         * read past the grid, which is reported instead */
        if (isx >= c->nx || isz >= c->nz) return 2;
        float vss[2][2];
        for (int k = 1; k <= 2; k++)
            for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = c->vel[IX(c, isz - 1 + l, isx - 1 + k)];
        float drx = (scx - c->gox) - (float)(isx - 1) * c->dnx;
        float drz = (scz - c->goz) - (float)(isz - 1) * c->dnz;
        float vels = bilinear(vss, drx, drz, c->dnx, c->dnz);
        for (int k = 1; k <= 2; k++)
            for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = c->vel[IX(c, irz - 1 + l, irx - 1 + k)];
        drx = (rcx - c->gox) - (float)(irx - 1) * c->dnx;
        drz = (rcz - c->goz) - (float)(irz - 1) * c->dnz;
        float velr = bilinear(vss, drx, drz, c->dnx, c->dnz);
        trr = 2.0f * sred / (vels + velr);
    } else {
        float drx = (rcx - c->gox) - (float)(irx - 1) * c->dnx;
        float drz = (rcz - c->goz) - (float)(irz - 1) * c->dnz;
        trr = 0.0f;
        for (int k = 1; k <= 2; k++)
            for (int l = 1; l <= 2; l++) {
                float produ = (1.0f - fabsf(((float)(l - 1) * c->dnz - drz) / c->dnz)) *
                              (1.0f - fabsf(((float)(k - 1) * c->dnx - drx) / c->dnx));
                trr = trr + c->t[IX(c, irz - 1 + l, irx - 1 + k)] * produ;
            }
    }
    *out = trr;
    return 0;
}

/* One job.  pv: the period's (nvz+2)*(nvx+2) phase velocities (column index
 * fastest along x).  Outputs: ttn (nnz*nnx), optionally ttnr/nstsr (refined
 * field and status after the refined pass, nrz*nrx, with *nrx_out/*nrz_out),
 * times[nrc].  Returns 0, or 1 (source outside), 2 (band over maxbt), 3
 * (coarse nodes left unreached, see below), 4 (source on the last node in the
 * near-source branch), 5 (receiver outside). */
int orc_fmm_job(const orc_fmm_desc *d, const double *pv, float scx, float scz, int nrc, const float *rcx,
                const float *rcz, float *ttn, int *nrx_out, int *nrz_out, float *ttnr, int *nstsr,
                float *times) {
    float gox, goz, dnx, dnz, dvx, dvz;
    coarse_grid(d, &gox, &goz, &dnx, &dnz, &dvx, &dvz);
    int nnx = (d->nvx - 1) * d->gdx + 1, nnz = (d->nvz - 1) * d->gdz + 1;
    size_t nn = (size_t)nnx * nnz;
    int maxbt = (int)lroundf(d->snb * (float)nnx * (float)nnz);
    float *veln = malloc(nn * sizeof(float));
    orc_fmm_gridder(d, pv, veln);
    size_t nc = (size_t)(d->nvx + 2) * (d->nvz + 2);
    float *velv = malloc(nc * sizeof(float));
    for (size_t c = 0; c < nc; c++) velv[c] = (float)pv[c];

    grid_t c = {0};
    c.nx = nnx;
    c.nz = nnz;
    c.gox = gox;
    c.goz = goz;
    c.dnx = dnx;
    c.dnz = dnz;
    c.vel = veln;
    c.t = ttn;
    c.st = malloc(nn * sizeof(int));
    c.earth = d->earth;
    c.fom = d->fom;
    int rc = 0;
    if (nrx_out) *nrx_out = 0;
    if (nrz_out) *nrz_out = 0;
    if (d->asgr == 1) {
        int isx = (int)((scx - gox) / dnx) + 1, isz = (int)((scz - goz) / dnz) + 1;
        if (isx < 1 || isx > nnx || isz < 1 || isz > nnz) {
            rc = 1;
            goto done;
        }
        if (isx == nnx) isx--;
        if (isz == nnz) isz--;
        int vnl = isx - d->sgs < 1 ? 1 : isx - d->sgs;
        int vnr = isx + d->sgs > nnx ? nnx : isx + d->sgs;
        int vnt = isz - d->sgs < 1 ? 1 : isz - d->sgs;
        int vnb = isz + d->sgs > nnz ? nnz : isz + d->sgs;
        int nrx = (vnr - vnl) * d->sgdl + 1, nrz = (vnb - vnt) * d->sgdl + 1;
        size_t nr = (size_t)nrx * nrz;
        int ldx = nrx > nnx ? nrx : nnx, ldz = nrz > nnz ? nrz : nnz;
        grid_t r = {0};
        r.nx = nrx;
        r.nz = nrz;
        r.dnx = dvx / (float)(d->gdx * d->sgdl);
        r.dnz = dvz / (float)(d->gdz * d->sgdl);
        r.gox = gox + dnx * (float)(vnl - 1);
        r.goz = goz + dnz * (float)(vnt - 1);
        float *velr = malloc(nr * sizeof(float));
        float *tr = malloc(nr * sizeof(float));
        r.vel = velr;
        r.t = tr;
        r.st = malloc(nr * sizeof(int));
        r.earth = d->earth;
        r.fom = d->fom;
        r.vnl = vnl;
        r.vnr = vnr;
        r.vnt = vnt;
        r.vnb = vnb;
        /* the tree is reallocated to the larger of the two grids */
        r.maxbt = (nrx > nnx || nrz > nnz) ? (int)lroundf(d->snb * (float)ldx * (float)ldz) : maxbt;
        r.bx = malloc(((size_t)r.maxbt + 2) * sizeof(int16_t));
        r.bz = malloc(((size_t)r.maxbt + 2) * sizeof(int16_t));
        for (size_t q = 0; q < nr; q++) tr[q] = 0.0f;
        bsplrefine(d, velv, vnl, vnr, vnt, vnb, nrx, nrz, velr);
        rc = travel(&r, scx, scz, 1);
        if (nrx_out) *nrx_out = nrx;
        if (nrz_out) *nrz_out = nrz;
        if (ttnr) memcpy(ttnr, tr, nr * sizeof(float));
        if (nstsr) memcpy(nstsr, r.st, nr * sizeof(int));
        if (rc == 0) {
            /* map the refined pass onto the coarse grid, CalSurfGAniso_Joint.f90:606-630 */
            for (size_t q = 0; q < nn; q++) c.st[q] = -1;
            for (int k = 1; k <= nrz; k += d->sgdl) {
                int i1 = vnt + (k - 1) / d->sgdl;
                for (int l = 1; l <= nrx; l += d->sgdl) {
                    int i2 = vnl + (l - 1) / d->sgdl;
                    int s = r.st[IX(&r, k, l)];
                    c.st[IX(&c, i1, i2)] = s;
                    if (s >= 0) c.t[IX(&c, i1, i2)] = r.t[IX(&r, k, l)];
                }
            }
            /* alive nodes next to far nodes rejoin the band, :663-683 */
            for (int k = 1; k <= nnx; k++)
                for (int l = 1; l <= nnz; l++) {
                    if (c.st[IX(&c, l, k)] != 0) continue;
                    if (l - 1 >= 1 && c.st[IX(&c, l - 1, k)] == -1) c.st[IX(&c, l, k)] = 1;
                    if (l + 1 <= nnz && c.st[IX(&c, l + 1, k)] == -1) c.st[IX(&c, l, k)] = 1;
                    if (k - 1 >= 1 && c.st[IX(&c, l, k - 1)] == -1) c.st[IX(&c, l, k)] = 1;
                    if (k + 1 <= nnx && c.st[IX(&c, l, k + 1)] == -1) c.st[IX(&c, l, k)] = 1;
                }
            c.maxbt = r.maxbt;
            c.bx = r.bx;
            c.bz = r.bz;
            rc = travel(&c, scx, scz, 2);
            /* The refined pass stops when its band reaches the refined grid's
             * far x (z) edge unless vnr (vnb), a coarse index, equals the
             * refined node count (CalSurfG.f90:503-523); for a source close to
             * the model's south or east edge that happens before the band
             * reaches any coarse node, the coarse band starts empty and the
             * reference leaves ttn unset.  Reported as 3. */
            if (rc == 0)
                for (size_t q = 0; q < nn; q++)
                    if (c.st[q] != 0) {
                        rc = 3;
                        break;
                    }
        }
        free(velr);
        free(tr);
        free(r.st);
        free(r.bx);
        free(r.bz);
        c.bx = c.bz = NULL;
    } else {
        c.maxbt = maxbt;
        c.bx = malloc(((size_t)maxbt + 2) * sizeof(int16_t));
        c.bz = malloc(((size_t)maxbt + 2) * sizeof(int16_t));
        rc = travel(&c, scx, scz, 0);
        free(c.bx);
        free(c.bz);
    }
    if (rc == 0)
        for (int q = 0; q < nrc; q++) {
            int e = srtimes(&c, scx, scz, rcx[q], rcz[q], &times[q]);
            if (e) {
                rc = e == 1 ? 5 : 4;
                break;
            }
        }
done:
    free(c.st);
    free(veln);
    free(velv);
    return rc;
}

/* Many jobs, OpenMP over jobs (the CPU baseline).  Job q uses period
 * job_k[q] of pv (ncol x K, column fastest), source (scx[q], scz[q]) and
 * receivers rec_off[q]..rec_off[q+1]-1.  ttn may be NULL.  Returns the number
 * of failed jobs; status[q] holds each job's code. */
int orc_fmm_jobs(const orc_fmm_desc *d, const double *pv, int njob, const int *job_k, const float *scx,
                 const float *scz, const int64_t *rec_off, const float *rcx, const float *rcz, float *times,
                 float *ttn, int *status, int nthreads) {
    int nnx = (d->nvx - 1) * d->gdx + 1, nnz = (d->nvz - 1) * d->gdz + 1;
    size_t nn = (size_t)nnx * nnz, nc = (size_t)(d->nvx + 2) * (d->nvz + 2);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : bad)
    for (int q = 0; q < njob; q++) {
        float *t = ttn ? ttn + (size_t)q * nn : malloc(nn * sizeof(float));
        int64_t r0 = rec_off[q];
        int e = orc_fmm_job(d, pv + (size_t)job_k[q] * nc, scx[q], scz[q], (int)(rec_off[q + 1] - r0), rcx + r0,
                            rcz + r0, t, NULL, NULL, NULL, NULL, times + r0);
        status[q] = e;
        bad += e != 0;
        if (e) /* a failed job's times are NaN, as the product writes them */
            for (int64_t r = r0; r < rec_off[q + 1]; r++) times[r] = NAN;
        if (!ttn) free(t);
    }
    return bad;
}
