"""GPU parity of the traveltime stage (dazi_dk_travel_*): bit-identical to
the oracle (oracle/fmm_port.c, itself bit-identical to the reference's
routines of CalSurfG.f90) and to the reference outputs committed in
tests/golden/ref_travel.npz: every receiver time, every job status and every
coarse traveltime field."""
import numpy as np
import pytest

from oracle import port as P
from dazi_current_release_amd import configs
from dazi_current_release_amd import travel as T

pytestmark = pytest.mark.gpu

F = np.float32


def run_dev(d, G, pv, want_ttn=True):
    import torch
    plan = T.TravelPlan(d, G["job_k"], G["scx"], G["scz"], G["rec_off"], G["rcx"], G["rcz"])
    pvd = torch.as_tensor(np.ascontiguousarray(pv, np.float64), device="cuda")
    out = plan.run(pvd, want_ttn=want_ttn)
    torch.cuda.synchronize()
    res = (out["times"].cpu().numpy(), out["status"].cpu().numpy(),
           out["ttn"].cpu().numpy().reshape(len(G["job_k"]), -1) if want_ttn else None)
    plan.close()
    return res


def oracle(d, G, pv, **kw):
    od = P.fmm_desc(d.nvx, d.nvz, d.goxd, d.gozd, d.dvxd, d.dvzd, gdx=d.gdx, gdz=d.gdz,
                    sgdl=d.sgdl, sgs=d.sgs, asgr=d.asgr, fom=d.fom, snb=d.snb, earth=d.earth)
    return P.fmm_jobs(od, pv, G["job_k"], G["scx"], G["scz"], G["rec_off"], G["rcx"], G["rcz"],
                      want_ttn=True)


def assert_same(gpu, orc):
    t, st, ttn = gpu
    ot, ottn, ost = orc
    assert np.array_equal(st, ost), (st, ost)
    assert np.array_equal(t, ot, equal_nan=True)
    good = ost == 0
    assert np.array_equal(ttn[good], ottn[good])


@pytest.mark.parametrize("name", ["mixed", "first", "plain", "small"])
def test_matches_committed_reference_outputs(golden, name):
    g = golden("ref_travel.npz")
    nx, ny, K, asgr, fom = (int(v) for v in g[f"{name}_case"])
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K, asgr=asgr, fom=fom)
    G = {k: g[f"{name}_{k}"] for k in ("job_k", "scx", "scz", "rec_off", "rcx", "rcz")}
    print(f"travel fixture {name}: {len(G['job_k'])} jobs", flush=True)
    times, st, ttn = run_dev(d, G, g[f"{name}_pv"])
    assert np.array_equal(st, g[f"{name}_rc"])
    assert np.array_equal(times, g[f"{name}_times"])
    assert np.array_equal(ttn[g[f"{name}_ttn_jobs"]], g[f"{name}_ttn"])


@pytest.mark.parametrize("slots", [0, 3])
def test_matches_oracle_with_round_robin_slots(slots):
    nx, ny, K = 30, 27, 4
    G = configs.travel_geometry(nx, ny, K, 7, seed=21)
    pv = configs.smooth_pv(nx, ny, K, seed=22)
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K, slots=slots)
    print(f"travel {nx}x{ny} K={K} jobs={len(G['job_k'])} slots={slots}", flush=True)
    assert_same(run_dev(d, G, pv), oracle(d, G, pv))


def test_plain4096_grid_matches_oracle():
    """The bench's grid: 64 x 64 columns, 306 x 306 propagation nodes."""
    nx = ny = 64
    K = 30
    G = configs.travel_geometry(nx, ny, K, 12, seed=5, periods_per_source=1)
    pv = configs.smooth_pv(nx, ny, K, seed=6)
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K)
    print(f"travel plain_4096 grid jobs={len(G['job_k'])}", flush=True)
    assert_same(run_dev(d, G, pv), oracle(d, G, pv))


def test_edge_sources_overflow_and_plan_reuse():
    nx, ny, K = 22, 26, 2
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K)
    pi = F(3.1415926535898)
    gox = (F(90.0) - F(29.0)) * pi / F(180.0)
    goz = F(98.0) * pi / F(180.0)
    dn = F(0.25) * pi / F(180.0) / F(5)
    nnz, nnx = T.grid_shape(d)
    # sources: north-west corner cell, a node, by the south-east edges (the
    # reference leaves that field unset: status 3), and on the last column
    sx = np.array([gox + F(0.3) * dn, gox + dn * F(7), gox + dn * F(nnx - 1.4),
                   gox + dn * F(nnx - 1), gox + dn * F(40.5)], F)
    sz = np.array([goz + F(0.6) * dn, goz + dn * F(5), goz + dn * F(nnz - 1.7),
                   goz + dn * F(2.5), goz + dn * F(50.2)], F)
    rx = np.concatenate([sx[[0]] + F(0.2) * dn, sx[[1]], sx, [gox + dn * F(3.3)]]).astype(F)
    rz = np.concatenate([sz[[0]] + F(0.1) * dn, sz[[1]] + F(0.4) * dn, sz, [goz + dn * 9]]).astype(F)
    # job q has receivers [q, q+3) of rx/rz (the last job none)
    off = np.array([0, 3, 6, 8, 8, 8], np.int64)
    G = dict(job_k=np.array([0, 1, 0, 1, 0], np.int32), scx=sx, scz=sz, rec_off=off,
             rcx=rx[:8], rcz=rz[:8])
    pv = configs.smooth_pv(nx, ny, K, seed=9)
    orc = oracle(d, G, pv)
    assert list(orc[2]) == [0, 0, 3, 3, 0]
    print("travel edge sources", flush=True)
    assert_same(run_dev(d, G, pv), orc)
    # same plan, new maps
    import torch
    plan = T.TravelPlan(d, G["job_k"], G["scx"], G["scz"], G["rec_off"], G["rcx"], G["rcz"])
    for seed in (31, 32):
        pv2 = configs.smooth_pv(nx, ny, K, seed=seed)
        out = plan.run(torch.as_tensor(pv2, device="cuda"), want_ttn=True)
        torch.cuda.synchronize()
        got = (out["times"].cpu().numpy(), out["status"].cpu().numpy(),
               out["ttn"].cpu().numpy().reshape(5, -1))
        assert_same(got, oracle(d, G, pv2))
    plan.close()
    # a band capacity too small for the grid: status 2 and NaN times
    d2 = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K, snb=0.004)
    got = run_dev(d2, G, pv)
    o2 = oracle(d2, G, pv)
    assert (o2[2] == 2).any()
    assert_same(got, o2)


def test_band_beyond_the_lds_tree_spills_and_matches_oracle():
    """A 486 x 486 propagation grid (100 x 100 columns) with sources near the
    centre: the narrow band grows past the 1,024 tree entries kept in LDS
    (about 1,270 at its largest), so the tree's spill path runs."""
    nx = ny = 100
    K = 2
    pi = F(3.1415926535898)
    gox = (F(90.0) - F(29.0)) * pi / F(180.0)
    goz = F(98.0) * pi / F(180.0)
    dn = F(0.25) * pi / F(180.0) / F(5)
    rng = np.random.default_rng(17)
    sx = np.array([gox + dn * F(228.4), gox + dn * F(240.7), gox + dn * F(251.2)], F)
    sz = np.array([goz + dn * F(252.7), goz + dn * F(231.3), goz + dn * F(244.9)], F)
    rx = (gox + dn * rng.uniform(5, 480, 30)).astype(F)
    rz = (goz + dn * rng.uniform(5, 480, 30)).astype(F)
    G = dict(job_k=np.array([0, 1, 1, 0, 1, 0], np.int32), scx=np.r_[sx, sx].astype(F),
             scz=np.r_[sz, sz[::-1]].astype(F), rec_off=np.arange(0, 31, 5, dtype=np.int64),
             rcx=rx, rcz=rz)
    pv = configs.smooth_pv(nx, ny, K, seed=4)
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K)
    print("travel spill grid 486 x 486, 6 jobs", flush=True)
    assert_same(run_dev(d, G, pv), oracle(d, G, pv))


def test_host_entry_point_matches_device_entry_point():
    nx, ny, K = 22, 26, 3
    G = configs.travel_geometry(nx, ny, K, 4, seed=2)
    pv = configs.smooth_pv(nx, ny, K, seed=3)
    d = T.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K)
    t, st, ttn = T.travel_times(d, G["job_k"], G["scx"], G["scz"], G["rec_off"], G["rcx"],
                                G["rcz"], pv.T, want_ttn=True)
    g = run_dev(d, G, pv)
    assert np.array_equal(t, g[0]) and np.array_equal(st, g[1])
    assert np.array_equal(ttn.reshape(len(G["job_k"]), -1), g[2])


def test_example_station_layout_matches_reference(golden):
    """All 1,469 records of the reference example's data file (20,877
    receivers, Yunnan grid 38 x 42) on the example model's phase velocities,
    against the reference's travel/srtimes outputs."""
    g = golden("yunnan_travel.npz")
    G = configs.stations_radians(g)
    d = T.desc(38, 42, 29.0, 98.0, 0.25, 0.25, kmax=g["pv"].shape[0])
    print(f"travel example layout: {len(G['job_k'])} jobs", flush=True)
    times, st, _ = run_dev(d, G, g["pv"].astype(np.float64), want_ttn=False)
    assert np.array_equal(st, g["rc"])
    assert np.array_equal(times, g["times"])
