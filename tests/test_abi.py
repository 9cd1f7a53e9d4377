"""CPU tests of the C ABI library: it loads and exports every declared entry point.

No compute call is made here (no GPU in this environment); argument checks that
return before any HIP call are exercised.
"""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT
from dazi_current_release_amd import _lib

HEADER = os.path.join(ROOT, "include", "dazi_dk.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dazi_dk_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert L.dazi_dk_abi_version() == 1


def test_struct_layouts_match_header():
    assert C.sizeof(_lib.Desc) == 40
    assert C.sizeof(_lib.Tune) == 24
    assert C.sizeof(_lib.Layered) == 8 + 8 + 8 * 21
    assert C.sizeof(_lib.GrowDesc) == 32
    assert C.sizeof(_lib.TravelDesc) == 64


def test_invalid_sizes_rejected_before_any_gpu_call():
    L = _lib.lib()
    d = _lib.Desc(0, 1, 13, 18, 2, 0, 64, 0, 0.5, 0)
    buf = np.zeros(16, np.float64)
    p = buf.ctypes.data_as(C.c_void_p)
    rc = L.dazi_dk_iso(C.byref(d), p, None, None, p, p, p, None, None, None)
    assert rc == _lib.DAZI_DK_EINVAL
    assert b"bad sizes" in L.dazi_dk_last_error()
    rc = L.dazi_dk_iso(C.byref(d), None, None, None, p, p, p, None, None, None)
    assert rc == _lib.DAZI_DK_EINVAL


def test_tune_validation():
    L = _lib.lib()
    saved = _lib.Tune()
    assert L.dazi_dk_get_tune(C.byref(saved)) == 0
    bad = _lib.make_tune(12, 4, 0, 64, 2e-6)
    assert L.dazi_dk_set_tune(C.byref(bad)) == _lib.DAZI_DK_EINVAL
    bad = _lib.make_tune(16, 3, 0, 64, 2e-6)      # 48 threads: not whole waves
    assert L.dazi_dk_set_tune(C.byref(bad)) == _lib.DAZI_DK_EINVAL
    auto = _lib.make_tune(0, 0, 0, 64, 2e-6)      # lanes by batch size
    assert L.dazi_dk_set_tune(C.byref(auto)) == 0
    good = _lib.make_tune(32, 8, 1, 64, 1e-6)
    assert L.dazi_dk_set_tune(C.byref(good)) == 0
    got = _lib.Tune()
    L.dazi_dk_get_tune(C.byref(got))
    assert (got.lanes_per_scan, got.cols_per_wg, got.mode) == (32, 8, 1)
    assert L.dazi_dk_set_tune(C.byref(saved)) == 0


def test_noalloc_size_limit_checked_before_any_gpu_call():
    """The allocation-free entry point rejects a batch beyond its static work
    memory with DAZI_DK_ELIMIT before touching the device."""
    L = _lib.lib()
    buf = np.zeros(64, np.float64)
    p = buf.ctypes.data_as(C.c_void_p).value
    b = _lib.Layered()
    b.ncol, b.max_layers, b.nsolves = 4, 8, 8
    for f in ("nl", "loff", "nk", "toff", "thk", "vp", "vs", "rho", "periods", "c", "status"):
        setattr(b, f, p)
    b.nlayers_total = _lib.NOALLOC_SCRATCH_BYTES // 48 + 1
    assert L.dazi_dk_layered_noalloc_dev(C.byref(b), None, None) == _lib.DAZI_DK_ELIMIT
    assert b"allocation-free" in L.dazi_dk_last_error()
    b.nlayers_total = 0
    assert L.dazi_dk_layered_noalloc_dev(C.byref(b), None, None) == _lib.DAZI_DK_EINVAL
    b.nlayers_total, b.dc_dvs = 32, p
    assert L.dazi_dk_layered_noalloc_dev(C.byref(b), None, None) == _lib.DAZI_DK_EINVAL
