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


if __name__ == "__main__":
    main()
