"""GPU test of the Fortran binding: depthkernel_hip / depthkernelTI_hip called
from a Fortran host with the reference's argument lists give the same outputs
as the Python mirror over the same C ABI, and the reference's own values."""
import numpy as np
import pytest

from conftest import load_golden
import fortran_io as FIO

pytestmark = pytest.mark.gpu


def test_fortran_host_matches_python_and_reference(tmp_path):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as dk
    f = load_golden("ref_plain_1024.npz")
    nz = f["vel"].shape[1]
    vel = np.asfortranarray(f["vel"].reshape(8, 4, nz, order="F"))
    dump, dump_ti = tmp_path / "dump", tmp_path / "dump_ti"
    r, fout = FIO.run(str(tmp_path), vel, f["depz"], f["periods"], float(f["minthk"]),
                      env=dict(DAZI_DEPTHKERNEL_DUMP_DIR=str(dump),
                               DAZI_DEPTHKERNEL_TI_DUMP_DIR=str(dump_ti)))
    assert r.returncode == 0, r.stderr
    assert "Depth kernel backend: hip" in r.stdout
    o = FIO.read_output(fout, 32, len(f["periods"]), nz)
    pv, svs, svp, srho = dk.run_depthkernel(vel, f["periods"], f["depz"], float(f["minthk"]))
    for name, want in (("pv", pv), ("sen_vs", svs), ("sen_vp", svp), ("sen_rho", srho)):
        assert np.array_equal(o[name], want), name
    assert np.array_equal(o["pv"] == 0, f["pv"] == 0)
    assert np.abs(o["pv"] - f["pv"]).max() <= 1e-5
    # reference dump files (CalSurfG.f90:115-140, depthkernelTI.f90:195-213)
    for name, fn in (("pv", "pv_depthkernel.bin"), ("sen_vs", "sen_vs_depthkernel.bin"),
                     ("sen_vp", "sen_vp_depthkernel.bin"), ("sen_rho", "sen_rho_depthkernel.bin")):
        assert np.array_equal(np.fromfile(dump / fn, np.float64), o[name].ravel(order="F"))
    assert np.array_equal(np.fromfile(dump_ti / "pv_ti.bin", np.float64),
                          o["ti_pv"].ravel(order="F"))
    assert np.array_equal(np.fromfile(dump_ti / "lsen_gsc_ti.bin", np.float32),
                          o["ti_lsen"].ravel(order="F"))
    tpv, tls = dk.run_depthkernelTI(vel, f["periods"], f["depz"], float(f["minthk"]))
    assert np.array_equal(o["ti_pv"], tpv) and np.array_equal(o["ti_lsen"], tls)
    n = f["ti_lsen"].shape[0]
    assert np.abs(o["ti_pv"][:n] - f["ti_pv"]).max() <= 1e-5
    d = o["ti_lsen"][:n] - f["ti_lsen"]
    assert np.abs(d).max() <= 1e-4
    assert (np.abs(d) / np.abs(f["ti_lsen"]).max(axis=-1, keepdims=True)).max() <= 1e-4


def test_fortran_dumps_ab_against_reference(tmp_path):
    """The reference's A/B workflow: its own FD dumps (the fixture, written in the
    dump layout) against the hip backend's dumps, compared with the A/B metrics and
    held to the reference's ISO bands (run_iso_depthkernel_ab.py:324-338)."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from dazi_current_release_amd import compare
    f = load_golden("ref_plain_1024.npz")
    nz, K = f["vel"].shape[1], len(f["periods"])
    vel = np.asfortranarray(f["vel"].reshape(8, 4, nz, order="F"))
    dump = tmp_path / "hip"
    r, _ = FIO.run(str(tmp_path), vel, f["depz"], f["periods"], float(f["minthk"]),
                   env=dict(DAZI_DEPTHKERNEL_DUMP_DIR=str(dump)))
    assert r.returncode == 0, r.stderr
    base = tmp_path / "ref"
    base.mkdir()
    for name, fn in (("pv", "pv_depthkernel.bin"), ("sen_vs", "sen_vs_depthkernel.bin"),
                     ("sen_vp", "sen_vp_depthkernel.bin"), ("sen_rho", "sen_rho_depthkernel.bin")):
        np.asfortranarray(f[name]).ravel(order="F").tofile(base / fn)
    rows = {x["array"]: x for x in compare.compare_dirs(base, dump, 8, 4, nz, K, vel)}
    assert rows["pv"]["max_abs"] <= 1e-5 and rows["pv"]["cross_abs_0.0001"] == 0
    for name, (l2, mx) in (("sen_vs", (1e-3, 2e-4)), ("sen_vp", (2e-3, 2e-4)),
                           ("sen_rho", (2e-3, 2e-4)), ("effective_dcdvs", (1e-3, 3e-4))):
        assert rows[name]["rel_l2"] <= l2 and rows[name]["max_abs"] <= mx, (name, rows[name])


def test_fortran_g_rows_match_the_device_path_and_the_oracle(tmp_path):
    """A Fortran G builder's loop: g_rows_hip_add_ray per ray (dense fdm maps,
    the reference's argument form) and one g_rows_hip_flush, against the
    Python host entry on the same stage outputs and against the oracle."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from oracle import port as P
    from dazi_current_release_amd import configs, grows
    f = load_golden("ref_plain_1024.npz")
    nz, K = f["vel"].shape[1], len(f["periods"])
    nx, ny = 8, 4
    vel = np.asfortranarray(f["vel"].reshape(nx, ny, nz, order="F"))
    R = configs.ray_workload(nx - 2, ny - 2, K, 3, seed=12, cell_km=5.0)
    nray = len(R["ray_k"])
    maps = [configs.ray_dense(R, r) for r in range(nray)]
    r, fout, fg = FIO.run(str(tmp_path), vel, f["depz"], f["periods"], float(f["minthk"]),
                          rays=(R["ray_k"] + 1, maps))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "G row assembly backend: hip" in r.stdout
    rw, iw, col = FIO.read_triplets(fg)
    o = FIO.read_output(fout, nx * ny, K, nz)
    # stage arrays in the library's (node, period, column) order
    st = lambda a: np.ascontiguousarray(np.moveaxis(a, [0, 1, 2], [2, 1, 0]))  # noqa: E731
    velz = np.ascontiguousarray(vel.reshape(nx * ny, nz, order="F").T)
    args = (velz, st(o["sen_vs"]), st(o["sen_vp"]), st(o["sen_rho"]), st(o["ti_lsen"]))
    want = grows.g_rows(R, nx - 2, ny - 2, *args)
    assert rw.size == want[0].size > 0
    for got, w in zip((rw, iw, col), want):
        assert np.array_equal(got, w)
    fdm, fdmc, fdms = (np.array([m[i] for m in maps]) for i in range(3))
    ref = P.g_rows(nx - 2, ny - 2, nz, R["ray_k"], fdm, fdmc, fdms, *args)
    for got, w in zip((rw, iw, col), ref):
        assert np.array_equal(got, w)


def test_fortran_surface_traveltimes_match_the_oracle(tmp_path):
    """The G builder's traveltime loop from a Fortran host: depthkernel_hip's
    phase velocities into surface_traveltimes_hip with the builder's geometry
    arrays, against the oracle's restatement of the reference's routines on
    the same phase velocities (bit-exact)."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from oracle import port as P
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_1024")
    nx, ny = W["nx"], W["ny"]
    nz, K = W["vel"].shape[-1], len(W["periods"])
    vel = np.asfortranarray(W["vel"])
    G = configs.travel_geometry(nx, ny, K, 4, seed=8)
    ftin, ftout = str(tmp_path / "trv.bin"), str(tmp_path / "trv_out.bin")
    FIO.write_travel(ftin, 29.0, 98.0, 0.25, 0.25, G, K)
    r, fout = FIO.run(str(tmp_path), vel, W["depz"], W["periods"], float(W["minthk"]),
                      extra=["-", "-", "-", "-", ftin, ftout])
    assert r.returncode == 0, r.stdout + r.stderr
    dsurf = FIO.read_travel(ftout)
    o = FIO.read_output(fout, nx * ny, K, nz)
    d = P.fmm_desc(nx - 2, ny - 2, 29.0, 98.0, 0.25, 0.25)
    want, _, st = P.fmm_jobs(d, np.ascontiguousarray(o["pv"].T), G["job_k"], G["scx"], G["scz"],
                             G["rec_off"], G["rcx"], G["rcz"])
    assert np.all(st == 0)
    assert dsurf.size == want.size and np.array_equal(dsurf, want)


def test_fortran_ti_timing_lines_and_failed_period_nan(tmp_path):
    """depthkernelTI_hip prints the reference's internal timing lines
    (depthkernelTI.f90:170-186) and, with DAZI_DEPTHKERNEL_TI_FAILED_NAN=1,
    leaves NaN in Lsen_Gsc exactly where the reference does: every interval of
    the periods without a phase velocity (its tregn96 at c = 0); without the
    flag those entries are 0."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    f = load_golden("ref_ti_failures.npz")
    ncol, nz = f["vel"].shape
    K = len(f["periods"])
    vel = np.asfortranarray(f["vel"].reshape(ncol, 1, nz, order="F"))
    outs = {}
    for nan in (False, True):
        env = dict(DAZI_DEPTHKERNEL_TI_FAILED_NAN="1") if nan else {}
        d = tmp_path / ("nan" if nan else "zero")
        d.mkdir()
        r, fout = FIO.run(str(d), vel, f["depz"], f["periods"], float(f["minthk"]), env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[nan] = FIO.read_output(fout, ncol, K, nz)
        lines = r.stdout.splitlines()
        assert "  depthkernelTI internal timing:" in lines
        for key in ("refine+surfdisp96 phase velocity = ", "tregn96 sensitivity kernel = ",
                    "splice refined kernel = ", "refined sublayers rmax min/max/avg = ",
                    "observed worker threads = "):
            assert any(line.startswith("    " + key) for line in lines), key
        rmax = [line for line in lines if "rmax min/max/avg" in line][0].split("=")[1].split("/")
        assert int(rmax[0]) == int(rmax[1]) > nz and float(rmax[2]) == int(rmax[0])
    failed = outs[True]["ti_pv"] == 0
    assert np.array_equal(failed, f["ti_pv"] == 0) and failed.any()
    got = outs[True]["ti_lsen"]
    assert np.array_equal(np.isnan(got), np.isnan(f["ti_lsen"]))
    assert np.all(np.isnan(got[failed]))
    assert np.all(outs[False]["ti_lsen"][failed] == 0)
    assert np.array_equal(got[~failed], outs[False]["ti_lsen"][~failed])
