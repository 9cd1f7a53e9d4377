"""A/B comparison of depth-kernel dumps (the reference's backend A/B workflow).

Reads two dump directories written by depthkernel / depthkernelTI
(DAZI_DEPTHKERNEL_DUMP_DIR: pv, sen_vs, sen_vp, sen_rho f64;
DAZI_DEPTHKERNEL_TI_DUMP_DIR: pv_ti f64, lsen_gsc_ti f32; all Fortran order,
layouts of code/src_inv_iso_joint/CalSurfG.f90:115-140 and
depthkernelTI.f90:195-213) and reports per array the metrics the reference's
comparator reports (code/depthkernel_lab/compare_depthkernel_dumps.py:36-57):
value ranges, max / mean / rms absolute difference, relative L2, and per
threshold the count of base values below it and of entries that cross it.
With the node Vs it also compares the G builder's effective dc/dVs
(CalSurfGAniso_Joint.f90:754-765).

With --nparpi it compares two G-triplet dumps instead (DAZI_G_TRIPLET_DUMP_DIR:
rw_triplets.bin f32, iw_triplets.bin and col_triplets.bin i32,
CalSurfG.f90:1572-1594), keyed by (iw, col), with the sets and metrics of the
reference's triplet comparator (code/depthkernel_lab/compare_g_triplet_dumps.py:
53-123): all entries, the common keys' differences (per Vs / Gc / Gs family),
and the keys present on one side only with their counts near ftol.

    python -m dazi_current_release_amd.compare --base A --test B --nx 64 --ny 64 --nz 25 --kmax 30 [--vel-bin vel.bin] [--out metrics.csv]
    python -m dazi_current_release_amd.compare --base A --test B --nparpi 92256 [--out metrics.csv]
"""
from __future__ import annotations

import argparse
import csv
import os

import numpy as np

LEVELS = (1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3)
ISO = {"pv": ("pv_depthkernel.bin", "ck"), "sen_vs": ("sen_vs_depthkernel.bin", "ckz"),
       "sen_vp": ("sen_vp_depthkernel.bin", "ckz"), "sen_rho": ("sen_rho_depthkernel.bin", "ckz")}
TI = {"pv_ti": ("pv_ti.bin", "ck", np.float64), "lsen_gsc": ("lsen_gsc_ti.bin", "ckj", np.float32)}


def read_dump(path, dtype, shape):
    a = np.fromfile(path, dtype=dtype)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{path}: {a.size} values, expected {int(np.prod(shape))}")
    return a.reshape(shape, order="F")


def metrics(name, base, test):
    base = np.asarray(base, np.float64)
    test = np.asarray(test, np.float64)
    d = test - base
    ad = np.abs(d)
    row = dict(array=name, shape="x".join(map(str, base.shape)), n=int(base.size),
               base_min=float(base.min()), base_max=float(base.max()),
               test_min=float(test.min()), test_max=float(test.max()),
               max_abs=float(ad.max()), mean_abs=float(ad.mean()),
               rms_abs=float(np.sqrt(np.mean(d * d))),
               rel_l2=float(np.linalg.norm(d) / max(np.linalg.norm(base), 1e-30)))
    for lv in LEVELS:
        row[f"base_abs_lt_{lv:g}"] = int(np.count_nonzero(np.abs(base) < lv))
        row[f"cross_abs_{lv:g}"] = int(np.count_nonzero((np.abs(base) <= lv) != (np.abs(test) <= lv)))
    return row


def brocher_coefficients(vs):
    """dVp/dVs and dRho/dVs of the Brocher relation in float32 (the G builder's
    coe_a, coe_rho)."""
    v = np.asarray(vs, np.float32)
    f = np.float32
    v2 = v * v
    v3 = v2 * v
    v4 = v3 * v
    a = f(2.0947) - f(0.8206) * f(2.0) * v + f(0.2683) * f(3.0) * v2 - f(0.0251) * f(4.0) * v3
    p = f(0.9409) + f(2.0947) * v - f(0.8206) * v2 + f(0.2683) * v3 - f(0.0251) * v4
    p2 = p * p
    p3 = p2 * p
    p4 = p3 * p
    r = a * (f(1.6612) - f(0.4721) * f(2.0) * p + f(0.0671) * f(3.0) * p2 -
             f(0.0043) * f(4.0) * p3 + f(0.000106) * f(5.0) * p4)
    return a.astype(np.float64), r.astype(np.float64)


def compare_dirs(base, test, nx, ny, nz, kmax, vel=None):
    """Metrics rows for every array present in both directories."""
    nxy = nx * ny
    shapes = {"ck": (nxy, kmax), "ckz": (nxy, kmax, nz), "ckj": (nxy, kmax, nz - 1)}
    rows, A, B = [], {}, {}
    for name, (fn, shp) in ISO.items():
        pa, pb = os.path.join(base, fn), os.path.join(test, fn)
        if os.path.exists(pa) and os.path.exists(pb):
            A[name] = read_dump(pa, np.float64, shapes[shp])
            B[name] = read_dump(pb, np.float64, shapes[shp])
            rows.append(metrics(name, A[name], B[name]))
    for name, (fn, shp, dt) in TI.items():
        pa, pb = os.path.join(base, fn), os.path.join(test, fn)
        if os.path.exists(pa) and os.path.exists(pb):
            rows.append(metrics(name, read_dump(pa, dt, shapes[shp]),
                                read_dump(pb, dt, shapes[shp])))
    if vel is not None and all(k in A for k in ("sen_vs", "sen_vp", "sen_rho")):
        v = np.asarray(vel, np.float32).reshape(nxy, nz, order="F")
        ca, cr = brocher_coefficients(v)
        eff = lambda D: D["sen_vs"] + D["sen_vp"] * ca[:, None, :] + D["sen_rho"] * cr[:, None, :]  # noqa
        rows.append(metrics("effective_dcdvs", eff(A), eff(B)))
    return rows


NEAR_FTOL = (1.0e-4, 1.1e-4, 1.5e-4, 2.0e-4, 5.0e-4, 1.0e-3)


def read_triplets(d):
    rw = np.fromfile(os.path.join(d, "rw_triplets.bin"), np.float32)
    iw = np.fromfile(os.path.join(d, "iw_triplets.bin"), np.int32)
    col = np.fromfile(os.path.join(d, "col_triplets.bin"), np.int32)
    if not rw.size == iw.size == col.size:
        raise ValueError(f"{d}: triplet arrays of sizes {rw.size}, {iw.size}, {col.size}")
    return iw, col, rw


def _abs_stats(v):
    a = np.abs(np.asarray(v, np.float64))
    if a.size == 0:
        return dict(n=0, min_abs=0.0, max_abs=0.0, mean_abs=0.0, rms=0.0)
    return dict(n=int(a.size), min_abs=float(a.min()), max_abs=float(a.max()),
                mean_abs=float(a.mean()), rms=float(np.sqrt((a * a).mean())))


def _families(col, nparpi):
    col = np.asarray(col)
    return dict(Vs=int(np.count_nonzero(col <= nparpi)),
                Gc=int(np.count_nonzero((col > nparpi) & (col <= 2 * nparpi))),
                Gs=int(np.count_nonzero(col > 2 * nparpi)))


def compare_triplets(base, test, nparpi):
    """Rows of the G-triplet A/B metrics for the triplets (iw, col, rw) of two
    runs (each a tuple, or a dump directory)."""
    if isinstance(base, str):
        base = read_triplets(base)
    if isinstance(test, str):
        test = read_triplets(test)
    (bi, bc, bv), (ti, tc, tv) = base, test
    key = lambda i, c: i.astype(np.int64) * (3 * nparpi + 1) + c.astype(np.int64)  # noqa: E731
    kb, kt = key(bi, bc), key(ti, tc)
    common, ib, it = np.intersect1d(kb, kt, return_indices=True)
    bonly = np.setdiff1d(np.arange(kb.size), ib)
    tonly = np.setdiff1d(np.arange(kt.size), it)
    d = tv[it].astype(np.float64) - bv[ib].astype(np.float64)
    cb = bv[ib].astype(np.float64)
    rel = lambda dd, bb: float(np.linalg.norm(dd) / max(np.linalg.norm(bb), 1.0e-30))  # noqa
    rows = [dict(set="base_all", **_abs_stats(bv), **_families(bc, nparpi)),
            dict(set="test_all", **_abs_stats(tv), **_families(tc, nparpi)),
            dict(set="common_diff", **_abs_stats(d), rel_l2=rel(d, cb))]
    cc = bc[ib]
    for fam, m in (("Vs", cc <= nparpi), ("Gc", (cc > nparpi) & (cc <= 2 * nparpi)),
                   ("Gs", cc > 2 * nparpi)):
        rows.append(dict(set="common_diff_" + fam, **_abs_stats(d[m]), rel_l2=rel(d[m], cb[m])))
    for name, idx, vals, cols in (("base_only", bonly, bv, bc), ("test_only", tonly, tv, tc)):
        r = dict(set=name, **_abs_stats(vals[idx]), **_families(cols[idx], nparpi))
        a = np.abs(vals[idx].astype(np.float64))
        for lev in NEAR_FTOL:
            r[f"abs_le_{lev:g}"] = int(np.count_nonzero(a <= lev))
        rows.append(r)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--base", required=True)
    ap.add_argument("--test", required=True)
    ap.add_argument("--nx", type=int)
    ap.add_argument("--ny", type=int)
    ap.add_argument("--nz", type=int)
    ap.add_argument("--kmax", type=int)
    ap.add_argument("--vel-bin", help="float32 node Vs (nx, ny, nz) Fortran order")
    ap.add_argument("--nparpi", type=int, help="G triplets: model columns per family")
    ap.add_argument("--out", help="CSV of all metrics")
    a = ap.parse_args(argv)
    if a.nparpi:
        rows = compare_triplets(a.base, a.test, a.nparpi)
        _write_csv(a.out, rows)
        for r in rows:
            print(f"{r['set']}: n={r['n']} max_abs={r['max_abs']:.6g} "
                  + (f"rel_l2={r['rel_l2']:.6g}" if "rel_l2" in r else ""))
        return rows
    if None in (a.nx, a.ny, a.nz, a.kmax):
        ap.error("--nx, --ny, --nz and --kmax are required for depth-kernel dumps")
    vel = None
    if a.vel_bin:
        vel = np.fromfile(a.vel_bin, np.float32).reshape((a.nx, a.ny, a.nz), order="F")
    rows = compare_dirs(a.base, a.test, a.nx, a.ny, a.nz, a.kmax, vel)
    _write_csv(a.out, rows)
    for r in rows:
        print(f"{r['array']}: max_abs={r['max_abs']:.6g} rms_abs={r['rms_abs']:.6g} "
              f"rel_l2={r['rel_l2']:.6g} cross_abs_1e-4={r['cross_abs_0.0001']}")
    return rows


def _write_csv(path, rows):
    if not path:
        return
    keys = sorted({k for r in rows for k in r})
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=keys)
        w.writeheader()
        w.writerows(rows)


if __name__ == "__main__":
    main()
