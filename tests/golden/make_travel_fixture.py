"""Writes tests/golden/ref_travel.npz: the reference's own traveltime routines
(gridder, bsplrefine, travel, srtimes of CalSurfG.f90, compiled in place by
oracle/Makefile into oracle/_ref/libdazi_ref.so and driven per (period, source)
job by oracle/fmm_ref_driver.f90) on seeded synthetic phase-velocity maps and
station layouts.  Only data values are stored: the inputs and the reference's
outputs (receiver times and status of every job, the coarse traveltime field of
every 6th job).

Cases (all single precision as the reference):
  * "mixed":  22 x 26 grid, 3 periods, 6 stations (18 jobs), asgr=1, fom=1
  * "first":  same geometry, fom=0 (first-order updates)
  * "plain":  same geometry, asgr=0 (no source-grid refinement)
  * "small":  11 x 9 grid (its refined source grid is larger than the coarse
              grid, so the reference reallocates its tree), 2 periods, 4 stations

Also writes tests/golden/yunnan_travel.npz: the station layout of the
reference's example (examples/test4_Yunnan/voro/China_YN_Rayleigh_RS_5-40s.dat:
1,469 (period, source) records with 20,877 receivers on the 38 x 42 grid of its
para.in, 0.25 degree cells from 29 N, 98 E), the oracle's phase velocities of the
example's starting model (yunnan_golden.npz's MOD, float32-rounded as pvRc holds
them), and the reference's receiver times for every record on them.

Usage: python tests/golden/make_travel_fixture.py  (needs `make -C oracle ref`)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dazi_current_release_amd import configs  # noqa: E402
from oracle import port as P  # noqa: E402

CASES = {
    "mixed": dict(nx=22, ny=26, K=3, nsta=6, asgr=1, fom=1, seed=3),
    "first": dict(nx=22, ny=26, K=3, nsta=6, asgr=1, fom=0, seed=3),
    "plain": dict(nx=22, ny=26, K=3, nsta=6, asgr=0, fom=1, seed=3),
    "small": dict(nx=11, ny=9, K=2, nsta=4, asgr=1, fom=1, seed=4),
}


def case_inputs(c):
    G = configs.travel_geometry(c["nx"], c["ny"], c["K"], c["nsta"], seed=c["seed"])
    pv = configs.smooth_pv(c["nx"], c["ny"], c["K"], seed=c["seed"] + 10)
    d = P.fmm_desc(c["nx"] - 2, c["ny"] - 2, 29.0, 98.0, 0.25, 0.25, asgr=c["asgr"],
                   fom=c["fom"])
    return G, pv, d


YUNNAN = "/root/reference/examples/test4_Yunnan/voro/China_YN_Rayleigh_RS_5-40s.dat"


def yunnan_stations():
    """Source headers '# lat lon period wavetype veltype' and receiver rows
    'lat lon velocity' in file order (Main_Jt.f90:387-437 reads them the same
    way); degrees as float32."""
    src, rec, off = [], [], [0]
    with open(YUNNAN) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if line.startswith("#"):
                src.append((float(t[1]), float(t[2]), int(t[3])))
                off.append(off[-1])
            else:
                rec.append((float(t[0]), float(t[1])))
                off[-1] += 1
    src = np.array(src)
    rec = np.array(rec)
    return dict(src_lat=src[:, 0].astype(np.float32), src_lon=src[:, 1].astype(np.float32),
                period=src[:, 2].astype(np.int32), rec_off=np.array(off, np.int64),
                rec_lat=rec[:, 0].astype(np.float32), rec_lon=rec[:, 1].astype(np.float32))


def yunnan():
    g = np.load(os.path.join(HERE, "yunnan_golden.npz"), allow_pickle=False)
    vel, depz, periods = g["vel"], g["depz"], g["periods"]
    nx, ny, nz = vel.shape
    pv = P.depthkernel(vel, periods, depz, float(g["minthk"]), lib="port", kernels=False,
                       nthreads=8)                                  # (ncol, K)
    pv = np.asarray(pv, np.float64)
    pv32 = pv.astype(np.float32)
    assert np.array_equal(pv32.astype(np.float64), pv)
    S = yunnan_stations()
    G = configs.stations_radians(S)
    d = P.fmm_desc(nx - 2, ny - 2, 29.0, 98.0, 0.25, 0.25)
    pvT = np.ascontiguousarray(pv.T)
    times, rc = [], []
    for q in range(len(G["job_k"])):
        a, b = G["rec_off"][q], G["rec_off"][q + 1]
        r = P.fmm_job(d, pvT[G["job_k"][q]], G["scx"][q], G["scz"][q], G["rcx"][a:b],
                      G["rcz"][a:b], lib="ref")
        times.append(r["times"])
        rc.append(r["rc"])
    np.savez_compressed(os.path.join(HERE, "yunnan_travel.npz"), pv=pv32.T.copy(), **S,
                        times=np.concatenate(times).astype(np.float32), rc=np.array(rc, np.int32))
    print("wrote yunnan_travel.npz:", len(rc), "records,", int(G["rec_off"][-1]), "receivers")


def main():
    if not P.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref")
    out = {}
    for name, c in CASES.items():
        G, pv, d = case_inputs(c)
        times, ttn, rc = [], [], []
        for q in range(len(G["job_k"])):
            a, b = G["rec_off"][q], G["rec_off"][q + 1]
            r = P.fmm_job(d, pv[G["job_k"][q]], G["scx"][q], G["scz"][q], G["rcx"][a:b],
                          G["rcz"][a:b], lib="ref")
            times.append(r["times"])
            ttn.append(r["ttn"])
            rc.append(r["rc"])
        for k, v in G.items():
            out[f"{name}_{k}"] = v
        out[f"{name}_pv"] = pv
        out[f"{name}_times"] = np.concatenate(times).astype(np.float32)
        keep = np.arange(0, len(ttn), 6, dtype=np.int32)  # coarse fields of every 6th job
        out[f"{name}_ttn_jobs"] = keep
        out[f"{name}_ttn"] = np.stack([ttn[q] for q in keep]).astype(np.float32)
        out[f"{name}_rc"] = np.array(rc, np.int32)
        out[f"{name}_case"] = np.array([c["nx"], c["ny"], c["K"], c["asgr"], c["fom"]], np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_travel.npz"), **out)
    print("wrote ref_travel.npz:", {k: v.shape for k, v in out.items() if k.endswith("ttn")})
    yunnan()


if __name__ == "__main__":
    main()
