"""CPU tests of the G-row assembly's oracle and host logic.

The oracle (oracle.port.g_rows) restates the reference's row fill and
compression (CalSurfGAniso_Joint.f90:746-797).  No reference output of G exists
to compare with, so it is pinned by what fixes it independently:
* the Brocher row coefficients are the derivatives dVp/dVs and drho/dVs of the
  Brocher (2005) polynomials (central differences in float64);
* the column layout, decoded with the reference's own decoding of a triplet
  column (block, depth, cell; diagnose_g_triplet_pattern.py:23-35), puts a
  single marked (column, period, depth) entry where it belongs, with the gate
  on |fdm| and the |value| > ftol filter;
* row sums are linear in the path weights (sum of the Gc block =
  sum_cells fdmc * sum_z Lsen, up to the dropped |value| <= ftol entries);
* the triplets pass the reference's own diagnostics (Main_Jt.f90:1652-1767).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import port as P
from dazi_current_release_amd import _lib, configs
from dazi_current_release_amd.grows import check_triplets, dump_g_triplets


def test_row_coefficients_are_brocher_derivatives():
    v = np.linspace(1.0, 4.8, 97)
    vp = lambda x: np.polyval([-0.0251, 0.2683, -0.8206, 2.0947, 0.9409], x)  # noqa: E731
    rho = lambda p: np.polyval([0.000106, -0.0043, 0.0671, -0.4721, 1.6612, 0.0], p)  # noqa
    h = 1e-5
    da = (vp(v + h) - vp(v - h)) / (2 * h)
    dr = (rho(vp(v + h)) - rho(vp(v - h))) / (2 * h)
    a, r = P.brocher_row_coeffs(v.astype(np.float32))
    assert np.abs(a - da).max() <= 2e-6 * np.abs(da).max()
    # float32 evaluation with cancellation between the terms: ~1e-5 relative
    assert np.abs(r - dr).max() <= 2e-5 * np.abs(dr).max()


def stage_arrays(nvx, nvz, nz, K, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    ncol = (nvx + 2) * (nvz + 2)
    vel = rng.uniform(2.5, 4.5, (nz, ncol)).astype(np.float32)
    s = lambda: rng.standard_normal((nz, K, ncol)) * 1e-3 * scale  # noqa: E731
    lsen = (rng.standard_normal((nz - 1, K, ncol)) * 1e-3 * scale).astype(np.float32)
    return vel, s(), s(), s(), lsen


def decode(col, nparpi, cells):
    # diagnose_g_triplet_pattern.py:23-35: block and depth of a 1-based column
    local0 = col.astype(np.int64) - 1
    return local0 // nparpi, (local0 % nparpi) // cells, (local0 % nparpi) % cells


def test_oracle_places_a_marked_entry_and_applies_the_gates():
    nvx, nvz, nz, K = 5, 4, 6, 3
    cells, nd = nvx * nvz, nz - 1
    nparpi = cells * nd
    ncol = (nvx + 2) * (nvz + 2)
    vel = np.full((nz, ncol), 3.0, np.float32)
    svs = np.zeros((nz, K, ncol))
    svp = np.zeros_like(svs)
    srho = np.zeros_like(svs)
    lsen = np.zeros((nd, K, ncol), np.float32)
    jj, kk, k, z = 3, 2, 1, 4              # interior cell (jj, kk), period, depth
    post = jj * (nvx + 2) + kk
    svs[z, k, post] = 0.25
    lsen[z, k, post] = np.float32(0.5)
    fdm = np.zeros((1, nvz + 2, nvx + 2), np.float32)
    fdmc = np.zeros_like(fdm)
    fdms = np.zeros_like(fdm)
    fdm[0, jj, kk], fdmc[0, jj, kk], fdms[0, jj, kk] = 2.0, 0.75, -0.125
    # a gated cell (|fdm| < ftol) with large fdmc, and a cell whose entries are tiny
    fdm[0, 1, 1], fdmc[0, 1, 1] = 5e-5, 100.0
    lsen[z, k, 1 * (nvx + 2) + 1] = 1.0
    fdm[0, 4, 5] = 1.0
    svs[0, k, 4 * (nvx + 2) + 5] = 5e-5
    rw, iw, col = P.g_rows(nvx, nvz, nz, [k], fdm, fdmc, fdms, vel, svs, svp, srho, lsen,
                           row_base=7)
    assert list(iw) == [7, 7, 7]
    b, d, c = decode(col, nparpi, cells)
    assert list(b) == [0, 1, 2] and list(d) == [z] * 3 and list(c) == [(jj - 1) * nvx + kk - 1] * 3
    assert rw.tolist() == [0.5, 0.375, -0.0625]
    assert np.all(np.diff(col) > 0)


def test_oracle_rows_are_linear_in_the_path_weights():
    R = configs.ray_workload(6, 5, 3, 4, seed=3)
    nvx, nvz, nz, K = 6, 5, 7, 3
    vel, svs, svp, srho, lsen = stage_arrays(nvx, nvz, nz, K, seed=1, scale=100.0)
    dense = [configs.ray_dense(R, r) for r in range(len(R["ray_k"]))]
    fdm, fdmc, fdms = (np.array([m[i] for m in dense]) for i in range(3))
    rw, iw, col = P.g_rows(nvx, nvz, nz, R["ray_k"], fdm, fdmc, fdms, vel, svs, svp, srho, lsen)
    cells, nd = nvx * nvz, nz - 1
    nparpi = cells * nd
    info = check_triplets(iw, col, rw, len(R["ray_k"]), 3 * nparpi)
    assert info["breaks"] == 0 and info["sum_cnt"] == rw.size
    for r in range(len(R["ray_k"])):
        k = R["ray_k"][r]
        sel = (iw == r + 1) & (col > nparpi) & (col <= 2 * nparpi)
        s = slice(R["cell_off"][r], R["cell_off"][r + 1])
        c = R["cell"][s]
        post = (c // nvx + 1) * (nvx + 2) + c % nvx + 1
        want = (lsen[:, k, post].astype(np.float64) * R["fdmc"][s]).sum()
        dropped = int(np.count_nonzero(np.abs(fdm[r]) >= 1e-4)) * nd - int(sel.sum())
        assert abs(rw[sel].astype(np.float64).sum() - want) <= 1e-4 * (dropped + 1) + \
            1e-5 * abs(want)


def test_check_triplets_flags_what_the_reference_stops_on():
    iw = np.array([1, 1, 2, 2], np.int32)
    col = np.array([1, 5, 2, 9], np.int32)
    rw = np.ones(4, np.float32)
    info = check_triplets(iw, col, rw, 2, 9)
    assert info["breaks"] == 0 and info["cnt"].tolist() == [2, 2]
    assert check_triplets(iw, np.array([1, 5, 9, 2], np.int32), rw, 2, 9)["breaks"] == 1
    with pytest.raises(ValueError, match="row-contiguous"):
        check_triplets(np.array([1, 2, 1, 2], np.int32), col, rw, 2, 9)
    with pytest.raises(ValueError, match="invalid"):
        check_triplets(iw, np.array([1, 5, 2, 10], np.int32), rw, 2, 9)
    with pytest.raises(ValueError, match="invalid"):
        check_triplets(iw, col, np.array([1, np.nan, 1, 1], np.float32), 2, 9)


def test_dump_format_is_the_reference_stream_layout(tmp_path, monkeypatch):
    rw = np.array([0.5, -2.0], np.float32)
    iw = np.array([3, 3], np.int32)
    col = np.array([4, 11], np.int32)
    monkeypatch.setenv("DAZI_G_TRIPLET_DUMP_DIR", str(tmp_path / "g"))
    dump_g_triplets(rw, iw, col)
    d = tmp_path / "g"
    assert np.array_equal(np.fromfile(d / "rw_triplets.bin", np.float32), rw)
    assert np.array_equal(np.fromfile(d / "iw_triplets.bin", np.int32), iw)
    assert np.array_equal(np.fromfile(d / "col_triplets.bin", np.int32), col)
    assert os.path.getsize(d / "rw_triplets.bin") == 8


def test_host_entry_rejects_bad_rays_before_any_gpu_call():
    L = _lib.lib()
    d = _lib.GrowDesc(4, 3, 5, 2, 1, 1, 1e-4)
    k = np.array([0], np.int32)
    off = np.array([0, 2], np.int64)
    w = np.ones(2, np.float32)
    big = np.zeros(10000, np.float64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    nar = C.c_int64(0)
    for cell, msg in [(np.array([3, 3], np.int32), b"strictly increasing"),
                      (np.array([0, 12], np.int32), b"outside")]:
        rc = L.dazi_dk_grows(C.byref(d), p(k), p(off), p(cell), p(w), p(w), p(w), p(big), p(big),
                             p(big), p(big), p(big), 0, None, None, None, C.byref(nar))
        assert rc == _lib.DAZI_DK_EINVAL
        assert msg in L.dazi_dk_last_error()
    k2 = np.array([2], np.int32)
    rc = L.dazi_dk_grows(C.byref(d), p(k2), p(off), p(np.array([0, 1], np.int32)), p(w), p(w),
                         p(w), p(big), p(big), p(big), p(big), p(big), 0, None, None, None,
                         C.byref(nar))
    assert rc == _lib.DAZI_DK_EINVAL and b"period index" in L.dazi_dk_last_error()
    bad = _lib.GrowDesc(0, 3, 5, 2, 1, 1, 1e-4)
    assert L.dazi_dk_grows_dev(C.byref(bad), *([None] * 11), 0, None, None, None, None,
                               C.byref(nar), None) == _lib.DAZI_DK_EINVAL
