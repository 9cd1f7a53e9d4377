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


def run(tmpdir, vel, depz, periods, minthk, timeout=300, env=None):
    fin, fout = os.path.join(tmpdir, "in.bin"), os.path.join(tmpdir, "out.bin")
    write_input(fin, vel, depz, periods, minthk)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=timeout,
                       env=e)
    return r, fout
