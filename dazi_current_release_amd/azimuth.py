"""Period azimuthal-anisotropy map on the device: the forward 2-psi Rayleigh
anisotropy of every interior cell and period from the stage's Lsen_Gsc and the
Gc, Gs layer models (code/src_inv_iso_joint/FwdAzimuthalAniMap.f90:44-89,
called from Main_Jt.f90:1578), and the reference's text output
period_Azm_tomo.inv (FwdAzimuthalAniMap.f90:64-86).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .stage import _ptr, _stream_ptr, _tptr

KEYS = ("cos", "sin", "amp", "amp_rela", "angle")


def azimuth_map_dev(lsen, gcf, gsf, tRcV, nx, ny, out=None, stream=None):
    """lsen (nz-1, K, nx*ny) f32 (the stage's layout), gcf, gsf (nz-1, ny-2, nx-2)
    f32 (the reference's gcf(nx-2, ny-2, nz-1)), tRcV (K, (nx-2)*(ny-2)) f64, all
    device tensors.  Returns dict of (K, (nx-2)*(ny-2)) float32 tensors: cos, sin
    (cosRc, sinRc), amp (AzimAmp), amp_rela (AzimAmpRela), angle (AzimAng,
    degrees)."""
    import torch
    nd, K, ncol = lsen.shape
    if ncol != nx * ny:
        raise ValueError(f"lsen has {ncol} columns, the grid {nx * ny}")
    cells = (nx - 2) * (ny - 2)
    if out is None:
        out = {k: torch.empty((K, cells), dtype=torch.float32, device=lsen.device) for k in KEYS}
    d = _lib.AzimDesc(nx, ny, nd + 1, K)
    rc = _lib.lib().dazi_dk_azimuth_map_dev(
        C.byref(d), _tptr(lsen), _tptr(gcf), _tptr(gsf), _tptr(tRcV),
        *[_tptr(out[k]) for k in KEYS], _stream_ptr(stream))
    _lib.check(rc)
    return out


def azimuth_map(lsen, gcf, gsf, tRcV, nx, ny):
    """Host arrays in the layouts of azimuth_map_dev; returns numpy arrays."""
    lsen = np.ascontiguousarray(lsen, np.float32)
    nd, K, ncol = lsen.shape
    cells = (nx - 2) * (ny - 2)
    out = {k: np.empty((K, cells), np.float32) for k in KEYS}
    d = _lib.AzimDesc(nx, ny, nd + 1, K)
    rc = _lib.lib().dazi_dk_azimuth_map(
        C.byref(d), _ptr(lsen), _ptr(np.ascontiguousarray(gcf, np.float32)),
        _ptr(np.ascontiguousarray(gsf, np.float32)), _ptr(np.ascontiguousarray(tRcV, np.float64)),
        *[_ptr(out[k]) for k in KEYS])
    _lib.check(rc)
    return out


def azm_lines(m, tRcV, tRc, nx, ny, goxd, gozd, dvxd, dvzd):
    """The reference's period_Azm_tomo.inv lines (FwdAzimuthalAniMap.f90:64-86,
    format '(10f10.5)'): for each period, row jj and column ii of the interior
    grid, lat = gozd + (jj-1) dvzd, lon = goxd - (ii-1) dvxd (float32), period,
    isoC = real(tRcV), AzimAng, AzimAmpRela, AzimAmp, cosRc, sinRc."""
    f = np.float32
    nvx, nvz = nx - 2, ny - 2
    lines = []
    for tt in range(len(tRc)):
        for jj in range(1, nvz + 1):
            for ii in range(1, nvx + 1):
                c = (jj - 1) * nvx + (ii - 1)
                vals = (f(gozd) + f(jj - 1) * f(dvzd), f(goxd) - f(ii - 1) * f(dvxd), tRc[tt],
                        f(tRcV[tt, c]), m["angle"][tt, c], m["amp_rela"][tt, c],
                        m["amp"][tt, c], m["cos"][tt, c], m["sin"][tt, c])
                lines.append("".join("%10.5f" % float(v) for v in vals))
    return lines
