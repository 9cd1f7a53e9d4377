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

    python -m dazi_current_release_amd.compare --base A --test B --nx 64 --ny 64 --nz 25 --kmax 30 [--vel-bin vel.bin] [--out metrics.csv]
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--base", required=True)
    ap.add_argument("--test", required=True)
    ap.add_argument("--nx", type=int, required=True)
    ap.add_argument("--ny", type=int, required=True)
    ap.add_argument("--nz", type=int, required=True)
    ap.add_argument("--kmax", type=int, required=True)
    ap.add_argument("--vel-bin", help="float32 node Vs (nx, ny, nz) Fortran order")
    ap.add_argument("--out", help="CSV of all metrics")
    a = ap.parse_args(argv)
    vel = None
    if a.vel_bin:
        vel = np.fromfile(a.vel_bin, np.float32).reshape((a.nx, a.ny, a.nz), order="F")
    rows = compare_dirs(a.base, a.test, a.nx, a.ny, a.nz, a.kmax, vel)
    if a.out:
        keys = sorted({k for r in rows for k in r})
        with open(a.out, "w", newline="") as fh:
            w = csv.DictWriter(fh, fieldnames=keys)
            w.writeheader()
            w.writerows(rows)
    for r in rows:
        print(f"{r['array']}: max_abs={r['max_abs']:.6g} rms_abs={r['rms_abs']:.6g} "
              f"rel_l2={r['rel_l2']:.6g} cross_abs_1e-4={r['cross_abs_0.0001']}")
    return rows


if __name__ == "__main__":
    main()
