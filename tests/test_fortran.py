"""The Fortran binding (module dazi_dk, reference argument lists) builds and
links against the C ABI; without a GPU it stops with the ABI's no-device code."""
import os

import numpy as np
import pytest

from conftest import load_golden
import fortran_io as FIO


def _gpu_present():
    return os.path.exists("/dev/kfd")


def test_driver_built():
    assert os.access(FIO.DRIVER, os.X_OK), "run __graft_entry__.build() first"


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device error path")
def test_no_device_stops_with_abi_code(tmp_path):
    f = load_golden("ref_plain_1024.npz")
    vel = f["vel"][:4].reshape(2, 2, -1, order="F")
    r, _ = FIO.run(str(tmp_path), vel, f["depz"], f["periods"], float(f["minthk"]))
    assert r.returncode == 9304 % 256, (r.returncode, r.stderr)
    assert "Depth kernel backend: hip" in r.stdout
    assert "ERROR: depthkernel hip backend: dazi_dk error 9304" in r.stdout
