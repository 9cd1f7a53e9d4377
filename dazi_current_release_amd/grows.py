"""G-row assembly of the joint inversion on the device (the consumer of the
stage's outputs).

The reference builds the sparse G matrix ray by ray inside its source/period/
record loops (code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:511-799): a dense
row of 3*nparpi float32 entries per ray from the path weights fdm, fdmc, fdms
and the stage's sen_vs/sen_vp/sen_rho and Lsen_Gsc (:750-787), compressed to
(rw, iw, col) triplets with |row| > ftol (:789-797).  Here a batch of rays
(each given by its touched cells) is assembled by the HIP library in one call
(``dazi_dk_grows_dev``); the triplets come out in the reference's order.

Also here: the reference's triplet dump (DAZI_G_TRIPLET_DUMP_DIR,
CalSurfG.f90:1572-1594) and its triplet diagnostics (Main_Jt.f90:1652-1767).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .stage import _ptr, _stream_ptr, _tptr

FTOL = 1.0e-4  # CalSurfGAniso_Joint.f90:330


def _desc(nvx, nvz, nz, K, nray, row_base, ftol):
    return _lib.GrowDesc(int(nvx), int(nvz), int(nz), int(K), int(nray), int(row_base),
                         float(ftol))


def rays_to_device(R, device="cuda"):
    """Host ray description (ray_k, cell_off, cell, fdm, fdmc, fdms) -> tensors."""
    import torch
    keys = ("ray_k", "cell_off", "cell", "fdm", "fdmc", "fdms")
    return {k: torch.as_tensor(np.ascontiguousarray(R[k]), device=device) for k in keys}


def g_rows_dev(rays, nvx, nvz, vel, sen_vs, sen_vp, sen_rho, lsen, row_base=1, ftol=FTOL,
               cap=None, out=None, stream=None):
    """Assembles the rows of a ray batch on the device.

    rays: dict of device tensors ray_k (nray,) i32, cell_off (nray+1,) i64,
    cell, fdm, fdmc, fdms; vel (nz, ncol) f32, sen_* (nz, K, ncol) f64, lsen
    (nz-1, K, ncol) f32 with ncol = (nvx+2)*(nvz+2).  With cap None the entry
    count is found by a first call (one synchronisation) and the outputs are
    sized to it; with out given (dict rw, iw, col) the call only enqueues.
    Returns dict(rw, iw, col, row_nnz, nar (1,) i64) on the device; entries
    beyond cap are not written (nar holds the full count)."""
    import torch
    nz, K, ncol = sen_vs.shape
    if ncol != (nvx + 2) * (nvz + 2):
        raise ValueError(f"stage arrays have {ncol} columns, grid needs {(nvx + 2) * (nvz + 2)}")
    nray = rays["ray_k"].numel()
    dev = sen_vs.device
    d = _desc(nvx, nvz, nz, K, nray, row_base, ftol)
    nar = torch.zeros(1, dtype=torch.int64, device=dev)
    row_nnz = torch.empty(max(nray, 1), dtype=torch.int64, device=dev)
    L = _lib.lib()

    def call(c, rw, iw, col):
        rc = L.dazi_dk_grows_dev(
            C.byref(d), _tptr(rays["ray_k"]), _tptr(rays["cell_off"]), _tptr(rays["cell"]),
            _tptr(rays["fdm"]), _tptr(rays["fdmc"]), _tptr(rays["fdms"]), _tptr(vel),
            _tptr(sen_vs), _tptr(sen_vp), _tptr(sen_rho), _tptr(lsen), int(c), _tptr(rw),
            _tptr(iw), _tptr(col), _tptr(row_nnz), _tptr(nar), _stream_ptr(stream))
        _lib.check(rc)

    if out is None:
        if cap is None:
            call(0, None, None, None)
            if stream is not None:
                stream.synchronize()
            else:
                torch.cuda.current_stream(dev).synchronize()
            cap = int(nar.item())
        out = dict(rw=torch.empty(max(cap, 1), dtype=torch.float32, device=dev),
                   iw=torch.empty(max(cap, 1), dtype=torch.int32, device=dev),
                   col=torch.empty(max(cap, 1), dtype=torch.int32, device=dev))
        out["cap"] = cap
    call(out.get("cap", out["rw"].numel()), out["rw"], out["iw"], out["col"])
    res = dict(out)
    res["row_nnz"] = row_nnz[:nray]
    res["nar"] = nar
    return res


def g_rows(R, nvx, nvz, vel, sen_vs, sen_vp, sen_rho, lsen, row_base=1, ftol=FTOL):
    """Host arrays in, host triplets out (dazi_dk_grows).  R: the ray description
    (numpy); vel (nz, ncol) f32, sen_* (nz, K, ncol) f64, lsen (nz-1, K, ncol)
    f32.  Returns rw f32, iw i32, col i32 in the reference's order."""
    sen_vs = np.ascontiguousarray(sen_vs, np.float64)
    nz, K, ncol = sen_vs.shape
    nray = len(R["ray_k"])
    cap = 3 * (nz - 1) * int(R["cell_off"][-1]) if nray else 0
    rw = np.empty(max(cap, 1), np.float32)
    iw = np.empty(max(cap, 1), np.int32)
    col = np.empty(max(cap, 1), np.int32)
    nar = C.c_int64(0)
    a = {k: np.ascontiguousarray(R[k]) for k in ("ray_k", "cell_off", "cell", "fdm", "fdmc",
                                                 "fdms")}
    arrs = [np.ascontiguousarray(vel, np.float32), sen_vs,
            np.ascontiguousarray(sen_vp, np.float64), np.ascontiguousarray(sen_rho, np.float64),
            np.ascontiguousarray(lsen, np.float32)]
    d = _desc(nvx, nvz, nz, K, nray, row_base, ftol)
    rc = _lib.lib().dazi_dk_grows(C.byref(d), _ptr(a["ray_k"]), _ptr(a["cell_off"]),
                                  _ptr(a["cell"]), _ptr(a["fdm"]), _ptr(a["fdmc"]),
                                  _ptr(a["fdms"]), *[_ptr(x) for x in arrs], cap, _ptr(rw),
                                  _ptr(iw), _ptr(col), C.byref(nar))
    _lib.check(rc)
    n = nar.value
    out = rw[:n].copy(), iw[:n].copy(), col[:n].copy()
    dump_g_triplets(*out)
    return out


def dump_g_triplets(rw, iw, col, dump_dir=None):
    """Writes rw_triplets.bin (f32), iw_triplets.bin and col_triplets.bin (i32)
    as raw streams into dump_dir, or into $DAZI_G_TRIPLET_DUMP_DIR when that is
    set (CalSurfG.f90:1572-1594; an empty value writes nothing)."""
    if dump_dir is None:
        dump_dir = os.environ.get("DAZI_G_TRIPLET_DUMP_DIR")
        if dump_dir is None:
            return
    dump_dir = dump_dir.strip()
    if not dump_dir:
        return
    os.makedirs(dump_dir, exist_ok=True)
    np.asarray(rw, np.float32).tofile(os.path.join(dump_dir, "rw_triplets.bin"))
    np.asarray(iw, np.int32).tofile(os.path.join(dump_dir, "iw_triplets.bin"))
    np.asarray(col, np.int32).tofile(os.path.join(dump_dir, "col_triplets.bin"))


def check_triplets(iw, col, rw, dall, nmodel):
    """The reference's triplet diagnostics before the Voronoi projection
    (Main_Jt.f90:1652-1767): counts of rows outside 1..dall, columns outside
    1..nmodel and non-finite values, whether iw is nondecreasing, the number of
    order breaks (a row decrease, or a column decrease within a row), and per-
    row counts.  Raises ValueError where the reference stops."""
    iw = np.asarray(iw)
    col = np.asarray(col)
    rw = np.asarray(rw)
    bad_row = int(np.count_nonzero((iw < 1) | (iw > dall)))
    bad_col = int(np.count_nonzero((col < 1) | (col > nmodel)))
    bad_nan = int(np.count_nonzero(~np.isfinite(rw.astype(np.float64))))
    dr = np.diff(iw.astype(np.int64))
    dc = np.diff(col.astype(np.int64))
    nondecreasing = bool(np.all(dr >= 0))
    breaks = int(np.count_nonzero(dr < 0) + np.count_nonzero((dr == 0) & (dc < 0)))
    ok = (iw >= 1) & (iw <= dall)
    cnt = np.bincount(iw[ok], minlength=dall + 1)[1:]
    info = dict(nar=int(iw.size), dall=int(dall), nmodel=int(nmodel), bad_row_count=bad_row,
                bad_col_count=bad_col, nonfinite_rw_count=bad_nan,
                iw_nondecreasing=nondecreasing, breaks=breaks, sum_cnt=int(cnt.sum()),
                cnt=cnt)
    if bad_row or bad_col or bad_nan:
        raise ValueError("invalid sparse triplets detected before Voronoi projection")
    if not nondecreasing:
        raise ValueError("sparse triplets are not row-contiguous before Voronoi projection")
    return info
