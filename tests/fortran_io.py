"""Stream-file I/O of the example Fortran host dazi_current_release_amd/fortran/dk_fdriver."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dazi_current_release_amd", "fortran", "dk_fdriver")


def write_input(path, vel, depz, periods, minthk):
    nx, ny, nz = vel.shape
    with open(path, "wb") as fh:
        np.array([nx, ny, nz, len(periods)], np.int32).tofile(fh)
        np.float32(minthk).tofile(fh)
        np.asarray(vel, np.float32).ravel(order="F").tofile(fh)
        np.asarray(depz, np.float32).tofile(fh)
        np.asarray(periods, np.float64).tofile(fh)


def read_output(path, ncol, K, nz):
    raw = open(path, "rb").read()
    out, off = {}, 0
    for name, shape, dt in (("pv", (ncol, K), np.float64), ("sen_vs", (ncol, K, nz), np.float64),
                            ("sen_vp", (ncol, K, nz), np.float64),
                            ("sen_rho", (ncol, K, nz), np.float64),
                            ("ti_pv", (ncol, K), np.float64),
                            ("ti_lsen", (ncol, K, nz - 1), np.float32)):
        n = int(np.prod(shape))
        a = np.frombuffer(raw, dt, count=n, offset=off)
        off += n * np.dtype(dt).itemsize
        out[name] = a.reshape(shape, order="F")
    assert off == len(raw), "unexpected output size"
    return out


def write_rays(path, knumi, maps):
    """knumi (nray,) 1-based; maps: per ray (fdm, fdmc, fdms) arrays [jj, kk] of
    shape (ny, nx), written as the reference's fdm(0:nvz+1, 0:nvx+1)."""
    with open(path, "wb") as fh:
        np.array([len(knumi)], np.int32).tofile(fh)
        np.asarray(knumi, np.int32).tofile(fh)
        for m in maps:
            for a in m:
                np.asarray(a, np.float32).ravel(order="F").tofile(fh)


def read_triplets(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int64, count=1)[0])
    rw = np.frombuffer(raw, np.float32, count=n, offset=8)
    iw = np.frombuffer(raw, np.int32, count=n, offset=8 + 4 * n)
    col = np.frombuffer(raw, np.int32, count=n, offset=8 + 8 * n)
    assert len(raw) == 8 + 12 * n, "unexpected triplet file size"
    return rw, iw, col


def run(tmpdir, vel, depz, periods, minthk, timeout=300, env=None, rays=None, extra=None):
    """Runs the driver; rays = (knumi, maps) adds the G-row step, whose
    triplets are read back as a third return value; extra: further driver
    arguments as given."""
    fin, fout = os.path.join(tmpdir, "in.bin"), os.path.join(tmpdir, "out.bin")
    write_input(fin, vel, depz, periods, minthk)
    args = [DRIVER, fin, fout]
    fg = None
    if rays is not None:
        frays, fg = os.path.join(tmpdir, "rays.bin"), os.path.join(tmpdir, "g.bin")
        write_rays(frays, *rays)
        args += [frays, fg]
    args += list(extra or [])
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=e)
    if rays is not None:
        return r, fout, fg
    return r, fout


def write_travel(path, goxd, gozd, dvxd, dvzd, G, K):
    """The G builder's geometry arrays for surface_traveltimes_hip from job
    arrays G (configs.travel_geometry, the same number of sources and receivers
    at every period): nsrcsurf1(K), periods, nrc1 (nsrcsurf, K), scxf, sczf
    (nsrcsurf, K), rcxf, rczf (nrcf, nsrcsurf, K), Fortran order."""
    njob = len(G["job_k"])
    ns = njob // K
    nr = int(np.diff(G["rec_off"]).max())
    assert ns * K == njob and np.all(np.diff(G["rec_off"]) == nr)
    assert np.array_equal(G["job_k"], np.repeat(np.arange(K), ns))
    periods = (G["job_k"].reshape(K, ns).T + 1).astype(np.int32)
    with open(path, "wb") as fh:
        np.array([goxd, gozd, dvxd, dvzd], np.float32).tofile(fh)
        np.array([ns, nr, K], np.int32).tofile(fh)
        np.full(K, ns, np.int32).tofile(fh)
        np.asfortranarray(periods).ravel(order="F").tofile(fh)
        np.full(ns * K, nr, np.int32).tofile(fh)
        for k in ("scx", "scz"):
            np.asarray(G[k], np.float32).reshape(K, ns).T.ravel(order="F").tofile(fh)
        for k in ("rcx", "rcz"):
            a = np.asarray(G[k], np.float32).reshape(K, ns, nr)   # [k, s, r]
            np.transpose(a, (2, 1, 0)).ravel(order="F").tofile(fh)


def read_travel(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int32, count=1)[0])
    return np.frombuffer(raw, np.float32, count=n, offset=4)
