"""Synthetic column-model workloads of the depth-kernel stage (seeded, numpy PCG64).

node workloads (the reference's depthkernel input: node Vs on a depth grid):
  plain_1024   32x32 columns, 13 nodes 0:5:60 km, minthk 0.5 (13 refined layers),
               Vs = (3.0 + 1.6 z/60)(1 + U(-5%,5%)), Vp = 1.73 Vs, rho = Brocher(Vp),
               18 log-spaced periods 2-40 s.
  plain_4096   64x64 columns, 25 nodes 0:5:120 km, minthk 1 (49 refined layers),
               crust 3.2->3.8 km/s above a U[30,55] km Moho, mantle 4.3->4.6 km/s,
               +-4 % Gaussian-smoothed (sigma 2 nodes) perturbation, Vp/Vs ~ U[1.70,1.80]
               per column, rho = Brocher(Vp), 30 log-spaced periods 1-60 s.
layered workloads (ragged layered models, the batched solver's input):
  ragged_16384 16,384 columns, U{8..64} layers: 0.1-3 km sediment (Vs 0.8-2.0),
               mid-crustal low-velocity zone (-5..-15 %) in 30 % of the columns, Moho
               U[25,60] km; U{10..40} sorted periods per column drawn from 1-60 s;
               Vp and rho from Brocher.
  diff_4096    4,096 columns x 40 layers x 30 periods (1-60 s), for the VJP workload;
               c_obs from a +-3 % Vs checkerboard of 4x4-column cells.
Voronoi realizations (an inversion step's population of models):
  voronoi_100  100 realizations; each has 9 depth layers, per layer 100 Vs cells with
               uniform-random nuclei over the 64x64 grid (nearest-nucleus assignment)
               and 50 ANI cells; projected onto 64x64 columns x 25 nodes 0:5:120 km
               (node z takes its layer's cell value): 409,600 node columns x 30
               periods; vp, rho by Brocher (the depthkernel semantics).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def brocher_vp(vs):
    """Brocher (2005) Vs->Vp in float32 with the reference's operation order
    (DepthKernel1D.f90:187-200)."""
    v = np.asarray(vs, F)
    v2 = v * v
    v3 = v2 * v
    v4 = v3 * v
    return (((F(0.9409) + F(2.0947) * v) - F(0.8206) * v2) + F(0.2683) * v3) - F(0.0251) * v4


def brocher_rho(vp):
    """Brocher (2005) Vp->rho in float32 (DepthKernel1D.f90:196-197)."""
    p = np.asarray(vp, F)
    p2 = p * p
    p3 = p2 * p
    p4 = p3 * p
    p5 = p4 * p
    return ((((F(1.6612) * p - F(0.4721) * p2) + F(0.0671) * p3) - F(0.0043) * p4)
            + F(0.000106) * p5)


def _smooth(x, sigma):
    """Gaussian smoothing along the last axis (reflect boundary)."""
    r = int(np.ceil(4 * sigma))
    t = np.arange(-r, r + 1)
    w = np.exp(-0.5 * (t / sigma) ** 2)
    w /= w.sum()
    xp = np.concatenate([x[..., r:0:-1], x, x[..., -2:-r - 2:-1]], axis=-1)
    out = np.zeros_like(x)
    for i, wi in enumerate(w):
        out += wi * xp[..., i:i + x.shape[-1]]
    return out


def node_workload(name: str, seed: int | None = None, ncol: int | None = None):
    """Node-model workload: dict(vel (nx,ny,nz) f32, vp, rho (same) or None,
    depz (nz,) f32, periods (K,) f64, minthk float, nx, ny)."""
    if name == "plain_1024":
        rng = np.random.default_rng(1 if seed is None else seed)
        nx = ny = 32
        depz = np.arange(0, 61, 5, dtype=F)
        nz = depz.size
        base = 3.0 + 1.6 * depz.astype(np.float64) / 60.0
        pert = rng.uniform(-0.05, 0.05, size=(nx, ny, nz))
        vs = (base * (1.0 + pert)).astype(F)
        vp = (F(1.73) * vs).astype(F)
        rho = brocher_rho(vp)
        periods = np.geomspace(2.0, 40.0, 18)
        minthk = 0.5
    elif name == "plain_4096":
        rng = np.random.default_rng(2 if seed is None else seed)
        nx = ny = 64
        depz = np.arange(0, 121, 5, dtype=F)
        nz = depz.size
        z = depz.astype(np.float64)
        moho = rng.uniform(30.0, 55.0, size=(nx, ny, 1))
        crust = 3.2 + 0.6 * z / moho
        mantle = 4.3 + 0.3 * (z - moho) / (120.0 - moho)
        base = np.where(z < moho, crust, mantle)
        noise = _smooth(rng.standard_normal((nx, ny, nz)), 2.0)
        noise /= noise.std(axis=-1, keepdims=True) + 1e-12
        pert = np.clip(0.02 * noise, -0.04, 0.04)
        vs = (base * (1.0 + pert)).astype(F)
        ratio = rng.uniform(1.70, 1.80, size=(nx, ny, 1))
        vp = (vs * ratio).astype(F)
        rho = brocher_rho(vp)
        periods = np.geomspace(1.0, 60.0, 30)
        minthk = 1.0
    else:
        raise ValueError(f"unknown node workload {name!r}")
    if ncol is not None:
        # a column subset (first ncol columns in Fortran order) as an (ncol,1,nz) grid
        sel = lambda a: np.ascontiguousarray(  # noqa: E731
            a.reshape(-1, a.shape[-1], order="F")[:ncol][:, None, :])
        vs, vp, rho = sel(vs), sel(vp), sel(rho)
        nx, ny = ncol, 1
    return dict(vel=vs, vp=vp, rho=rho, depz=depz, periods=periods, minthk=minthk, nx=nx, ny=ny)


VORONOI_LAYER_TOPS = np.array([0.0, 5.0, 10.0, 20.0, 30.0, 45.0, 60.0, 80.0, 100.0])


def voronoi_workload(nreal: int = 100, seed: int | None = None, ncells: int = 100,
                     nani: int = 50):
    """dict(vel (nreal*4096, 1, 25) f32 node Vs (column j of realization r at
    r*4096 + j, j = ix + 64*iy), depz, periods, minthk, layer_of_node (25,),
    cell (nreal, 9, 4096) Vs-cell index per layer and column, ani_cell (nreal,
    9, 4096), cell_vs (nreal, 9, ncells) f32)."""
    rng = np.random.default_rng(5 if seed is None else seed)
    n = 64
    depz = np.arange(0, 121, 5, dtype=F)
    z = depz.astype(np.float64)
    layer_of_node = np.searchsorted(VORONOI_LAYER_TOPS, z, side="right") - 1
    nl = VORONOI_LAYER_TOPS.size
    gx, gy = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="xy")
    pts = np.stack([gx.ravel(), gy.ravel()], axis=1)  # column j = ix + 64 iy
    mid = np.append(0.5 * (VORONOI_LAYER_TOPS[1:] + VORONOI_LAYER_TOPS[:-1]), 110.0)
    base = np.where(mid < 35.0, 3.2 + 0.6 * mid / 35.0, 4.3 + 0.3 * (mid - 35.0) / 85.0)

    p2 = (pts ** 2).sum(1)[:, None]

    def assign(nuclei):
        d2 = p2 - 2.0 * (pts @ nuclei.T) + (nuclei ** 2).sum(1)[None, :]
        return np.argmin(d2, axis=1).astype(np.int32)

    cell = np.empty((nreal, nl, n * n), np.int32)
    ani = np.empty((nreal, nl, n * n), np.int32)
    cell_vs = np.empty((nreal, nl, ncells), F)
    vel = np.empty((nreal, n * n, z.size), F)
    for r in range(nreal):
        for l in range(nl):
            cell[r, l] = assign(rng.uniform(0.0, n, size=(ncells, 2)))
            ani[r, l] = assign(rng.uniform(0.0, n, size=(nani, 2)))
            cell_vs[r, l] = (base[l] * (1.0 + rng.uniform(-0.05, 0.05, ncells))).astype(F)
        vel[r] = cell_vs[r][layer_of_node[None, :], cell[r][layer_of_node].T]
    return dict(vel=vel.reshape(nreal * n * n, 1, z.size), depz=depz,
                periods=np.geomspace(1.0, 60.0, 30), minthk=1.0, layer_of_node=layer_of_node,
                cell=cell, ani_cell=ani, cell_vs=cell_vs, nx=nreal * n * n, ny=1)


def _layered_profile(rng, L, moho, sed_thk, sed_vs, lvz, bottom=150.0):
    """One layered column: thk (L,), vs (L,), half-space last."""
    n_rest = L - 2  # layers between the sediment and the half-space
    edges = np.linspace(sed_thk, bottom, n_rest + 1)
    thk = np.empty(L)
    vs = np.empty(L)
    thk[0] = sed_thk
    vs[0] = sed_vs
    thk[1:L - 1] = np.diff(edges)
    mid = 0.5 * (edges[:-1] + edges[1:])
    crust = 3.2 + 0.6 * mid / moho
    mantle = 4.3 + 0.3 * (mid - moho) / (bottom - moho)
    v = np.where(mid < moho, crust, mantle)
    if lvz is not None:
        top, bot, drop = lvz
        v = np.where((mid >= top) & (mid < bot), v * (1.0 - drop), v)
    vs[1:L - 1] = v
    thk[L - 1] = 0.0
    vs[L - 1] = 4.7
    return thk, vs


def layered_workload(name: str, seed: int | None = None, ncol: int | None = None):
    """Ragged layered workload: dict(nl, nk (ncol,) i32, thk, vp, vs, rho (sum nl,) f32,
    periods (sum nk,) f64)."""
    if name == "ragged_16384":
        rng = np.random.default_rng(3 if seed is None else seed)
        n = 16384 if ncol is None else ncol
        nl = rng.integers(8, 65, size=n).astype(np.int32)
        nk = rng.integers(10, 41, size=n).astype(np.int32)
        thk_l, vs_l, per_l = [], [], []
        for j in range(n):
            moho = rng.uniform(25.0, 60.0)
            sed_thk = rng.uniform(0.1, 3.0)
            sed_vs = rng.uniform(0.8, 2.0)
            lvz = None
            if rng.uniform() < 0.3:
                top = rng.uniform(10.0, 18.0)
                lvz = (top, top + rng.uniform(4.0, 8.0), rng.uniform(0.05, 0.15))
            t, v = _layered_profile(rng, int(nl[j]), moho, sed_thk, sed_vs, lvz)
            thk_l.append(t)
            vs_l.append(v)
            p = np.sort(np.exp(rng.uniform(np.log(1.0), np.log(60.0), size=int(nk[j]))))
            per_l.append(p)
        thk = np.concatenate(thk_l).astype(F)
        vs = np.concatenate(vs_l).astype(F)
        periods = np.concatenate(per_l)
    elif name == "diff_4096":
        rng = np.random.default_rng(4 if seed is None else seed)
        n = 4096 if ncol is None else ncol
        L, K = 40, 30
        nl = np.full(n, L, np.int32)
        nk = np.full(n, K, np.int32)
        thk_l, vs_l = [], []
        for j in range(n):
            moho = rng.uniform(30.0, 55.0)
            t, v = _layered_profile(rng, L, moho, 2.0, rng.uniform(2.4, 3.0), None,
                                    bottom=160.0)
            v[:-1] *= 1.0 + rng.uniform(-0.02, 0.02, size=L - 1)
            thk_l.append(t)
            vs_l.append(v)
        thk = np.concatenate(thk_l).astype(F)
        vs = np.concatenate(vs_l).astype(F)
        periods = np.tile(np.geomspace(1.0, 60.0, K), n)
    else:
        raise ValueError(f"unknown layered workload {name!r}")
    vp = brocher_vp(vs)
    rho = brocher_rho(vp)
    return dict(nl=nl, nk=nk, thk=thk, vp=vp.astype(F), vs=vs, rho=rho.astype(F),
                periods=periods.astype(np.float64))


def checkerboard_vs(vs, nl, ncol_side=64, cell=4, amp=0.03):
    """+-amp Vs checkerboard of cell x cell column blocks over a ncol_side-wide grid
    (column j at (j % side, j // side)); the half-space is left unchanged."""
    out = np.array(vs, F, copy=True)
    loff = np.concatenate([[0], np.cumsum(nl)[:-1]])
    for j in range(len(nl)):
        ii, jj = j % ncol_side, j // ncol_side
        sgn = 1.0 if ((ii // cell) + (jj // cell)) % 2 == 0 else -1.0
        o = loff[j]
        out[o:o + nl[j] - 1] *= F(1.0 + sgn * amp)
    return out


def ray_workload(nvx: int, nvz: int, K: int, rays_per_period: int, seed: int = 7,
                 cell_km: float = 25.0):
    """Synthetic ray paths for the G-row assembly: straight rays between uniform
    random points of the interior grid (cells (jj, kk), jj = 1..nvz, kk = 1..nvx,
    cell kk spanning x in [kk - 0.5, kk + 0.5)), rays_per_period per period,
    period-major as the reference's knumi loop.  fdm = path length in the cell
    (km, sampled every 1/40 cell), fdmc = fdm cos 2phi, fdms = fdm sin 2phi with
    phi the ray azimuth.  Returns dict(ray_k (nray,) i32, cell_off (nray+1,) i64,
    cell (n,) i32 = (jj-1)*nvx + kk-1 ascending within a ray, fdm, fdmc, fdms (n,)
    f32, nvx, nvz)."""
    rng = np.random.default_rng(seed)
    nray = K * rays_per_period
    ray_k = np.repeat(np.arange(K, dtype=np.int32), rays_per_period)
    lo = np.array([0.5, 0.5])
    hi = np.array([nvx + 0.5, nvz + 0.5])
    p0 = rng.uniform(lo, hi - 1e-6, size=(nray, 2))
    p1 = rng.uniform(lo, hi - 1e-6, size=(nray, 2))
    d = p1 - p0
    length = np.hypot(d[:, 0], d[:, 1])
    phi = np.arctan2(d[:, 1], d[:, 0])
    step = 1.0 / 40.0
    ns = np.maximum(np.ceil(length / step).astype(np.int64), 1)
    cells = nvx * nvz
    offs, cl, w = [0], [], []
    chunk = 4096
    for a in range(0, nray, chunk):
        b = min(a + chunk, nray)
        n = ns[a:b]
        rid = np.repeat(np.arange(a, b), n)
        t = (np.arange(n.sum()) - np.repeat(np.cumsum(n) - n, n) + 0.5) / np.repeat(n, n)
        x = p0[rid, 0] + t * d[rid, 0]
        z = p0[rid, 1] + t * d[rid, 1]
        kk = np.clip(np.floor(x + 0.5).astype(np.int64), 1, nvx)
        jj = np.clip(np.floor(z + 0.5).astype(np.int64), 1, nvz)
        key = rid * cells + (jj - 1) * nvx + (kk - 1)
        u, inv = np.unique(key, return_inverse=True)
        seg = (length[rid] / np.repeat(n, n)) * cell_km
        wsum = np.bincount(inv, weights=seg)
        r_of = u // cells
        cnt = np.bincount(r_of - a, minlength=b - a)
        offs.extend((offs[-1] + np.cumsum(cnt)).tolist())
        cl.append((u % cells).astype(np.int32))
        w.append(wsum)
    cell = np.concatenate(cl)
    fdm = np.concatenate(w)
    rr = np.repeat(np.arange(nray), np.diff(np.array(offs)))
    return dict(ray_k=ray_k, cell_off=np.array(offs, np.int64), cell=cell,
                fdm=fdm.astype(F), fdmc=(fdm * np.cos(2 * phi[rr])).astype(F),
                fdms=(fdm * np.sin(2 * phi[rr])).astype(F), nvx=nvx, nvz=nvz)


def ray_dense(R, r):
    """fdm, fdmc, fdms of ray r of a ray_workload as the reference's dense maps
    fdm(0:nvz+1, 0:nvx+1): three (nvz+2, nvx+2) float32 arrays indexed [jj, kk]."""
    nvx, nvz = R["nvx"], R["nvz"]
    s = slice(R["cell_off"][r], R["cell_off"][r + 1])
    c = R["cell"][s]
    out = []
    for key in ("fdm", "fdmc", "fdms"):
        m = np.zeros((nvz + 2, nvx + 2), F)
        m[c // nvx + 1, c % nvx + 1] = R[key][s]
        out.append(m)
    return out


def travel_geometry(nx: int, ny: int, K: int, nsta: int, goxd: float = 29.0, gozd: float = 98.0,
                    dvxd: float = 0.25, dvzd: float = 0.25, seed: int = 11,
                    periods_per_source: int | None = None):
    """Station layout for the traveltime stage on an nx x ny inversion grid
    (border included; the propagation grid spans its interior nodes from
    latitude goxd southward and longitude gozd eastward).  nsta stations at
    uniform random interior positions; every period uses every station as a
    source (or periods_per_source of them, rotating) with all other stations as
    its receivers, period-major as the reference's knumi/srcnum loops.
    Coordinates are colatitude and longitude in radians, float32, as the G
    builder's scxf/sczf.  Returns dict(job_k, scx, scz, rec_off, rcx, rcz)."""
    rng = np.random.default_rng(seed)
    pi = np.float32(3.1415926535898)
    lat = rng.uniform(goxd - dvxd * (nx - 3) * 0.98, goxd - dvxd * (nx - 3) * 0.02, nsta)
    lon = rng.uniform(gozd + dvzd * (ny - 3) * 0.02, gozd + dvzd * (ny - 3) * 0.98, nsta)
    sx = ((np.float32(90.0) - lat.astype(F)) * pi / np.float32(180.0)).astype(F)
    sz = (lon.astype(F) * pi / np.float32(180.0)).astype(F)
    job_k, scx, scz, off, rcx, rcz = [], [], [], [0], [], []
    for k in range(K):
        srcs = range(nsta) if periods_per_source is None else \
            [(k * periods_per_source + i) % nsta for i in range(periods_per_source)]
        for s in srcs:
            job_k.append(k)
            scx.append(sx[s])
            scz.append(sz[s])
            others = np.array([r for r in range(nsta) if r != s], np.int64)
            rcx.append(sx[others])
            rcz.append(sz[others])
            off.append(off[-1] + others.size)
    return dict(job_k=np.array(job_k, np.int32), scx=np.array(scx, F), scz=np.array(scz, F),
                rec_off=np.array(off, np.int64),
                rcx=np.concatenate(rcx).astype(F) if rcx else np.zeros(0, F),
                rcz=np.concatenate(rcz).astype(F) if rcz else np.zeros(0, F))


def smooth_pv(nx: int, ny: int, K: int, seed: int = 5, lo: float = 2.8, hi: float = 4.2):
    """Smooth random phase-velocity maps: (K, nx*ny) float64, the column index
    ii + jj*nx fastest (pvRc in Fortran order, transposed)."""
    rng = np.random.default_rng(seed)
    out = np.empty((K, nx * ny))
    for k in range(K):
        base = rng.standard_normal((ny, nx))
        sx, sy = min(3.0, (nx - 1) / 5.0), min(3.0, (ny - 1) / 5.0)
        m = _smooth(_smooth(base, sx).T, sy).T
        m = (m - m.min()) / max(m.max() - m.min(), 1e-12)
        out[k] = (lo + (hi - lo) * m).reshape(-1) + 0.02 * k
    return out


def stations_radians(S):
    """Station records in degrees (src_lat, src_lon, period (1-based index),
    rec_off, rec_lat, rec_lon: one source header per (period, source) record,
    as a DAzi data file lists them) to traveltime job arrays: colatitude and
    longitude in radians computed in float32 as the data reader does
    ((90 - lat)*pi/180, lon*pi/180 with pi = 3.1415926535898 single precision,
    Main_Jt.f90:410-431), job_k the 0-based period index."""
    pi = np.float32(3.1415926535898)
    c = lambda lat: ((np.float32(90.0) - np.asarray(lat, F)) * pi / np.float32(180.0)).astype(F)  # noqa: E731
    z = lambda lon: (np.asarray(lon, F) * pi / np.float32(180.0)).astype(F)  # noqa: E731
    return dict(job_k=(np.asarray(S["period"], np.int32) - 1), scx=c(S["src_lat"]),
                scz=z(S["src_lon"]), rec_off=np.asarray(S["rec_off"], np.int64),
                rcx=c(S["rec_lat"]), rcz=z(S["rec_lon"]))
