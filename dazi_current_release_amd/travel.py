"""First-arrival traveltimes on the device: the consumer of the stage's
phase-velocity maps pvRc.

The reference computes them source by source inside the joint G builder
(code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:511-725): gridder dices a
period's map onto the propagation grid, bsplrefine and travel march the first
arrivals on a refined grid around the source and then on the coarse grid, and
srtimes interpolates each receiver's time (routines in CalSurfG.f90).  Here
all (period, source) jobs of a G build run in one call of the HIP library
(``dazi_dk_travel_run_dev``), bit-identical to the reference, with pvRc read
where the stage left it on the device.

A ``TravelPlan`` holds the geometry (sources, receivers, the sin tables the
marching reads) and the per-job work memory on the device; it is built once
per station layout and run on every new pvRc.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .stage import _ptr, _stream_ptr, _tptr

# CalSurfGAniso_Joint.f90:375-382
GDX = 5
SGDL = 8
SGS = 8
EARTH = 6371.0
SNB = 0.5


def desc(nx, ny, goxd, gozd, dvxd, dvzd, kmax, gdx=GDX, gdz=GDX, sgdl=SGDL, sgs=SGS, asgr=1,
         fom=1, snb=SNB, earth=EARTH, slots=0):
    """Descriptor for an nx x ny inversion grid (its border included, as the
    stage's columns): nvx = nx-2, nvz = ny-2, origin (goxd latitude, gozd
    longitude) and spacing in degrees, kmax periods in pvRc."""
    return _lib.TravelDesc(int(nx) - 2, int(ny) - 2, int(gdx), int(gdz), int(sgdl), int(sgs),
                           int(asgr), int(fom), float(goxd), float(gozd), float(dvxd),
                           float(dvzd), float(snb), float(earth), int(kmax), int(slots))


def grid_shape(d):
    """(nnz, nnx) of the propagation grid."""
    return (d.nvz - 1) * d.gdz + 1, (d.nvx - 1) * d.gdx + 1


def _geometry(job_k, scx, scz, rec_off, rcx, rcz):
    g = dict(job_k=np.ascontiguousarray(job_k, np.int32),
             scx=np.ascontiguousarray(scx, np.float32), scz=np.ascontiguousarray(scz, np.float32),
             rec_off=np.ascontiguousarray(rec_off, np.int64),
             rcx=np.ascontiguousarray(rcx, np.float32), rcz=np.ascontiguousarray(rcz, np.float32))
    n = g["job_k"].size
    if g["scx"].size != n or g["scz"].size != n or g["rec_off"].size != n + 1:
        raise ValueError("job arrays disagree in length")
    nrec = int(g["rec_off"][-1])
    if g["rcx"].size != nrec or g["rcz"].size != nrec:
        raise ValueError(f"rec_off[-1] = {nrec} receivers, arrays hold {g['rcx'].size}")
    return g


class TravelPlan:
    """Geometry and work memory of the traveltime jobs on the current device.

    job q: period job_k[q] (0-based), source (scx[q], scz[q]) colatitude and
    longitude in radians, receivers rec_off[q] .. rec_off[q+1]-1 of rcx/rcz."""

    def __init__(self, d, job_k, scx, scz, rec_off, rcx, rcz):
        self.desc = d
        self.geo = _geometry(job_k, scx, scz, rec_off, rcx, rcz)
        self.njob = self.geo["job_k"].size
        self.nrec = int(self.geo["rec_off"][-1])
        self._h = C.c_void_p()
        g = self.geo
        rc = _lib.lib().dazi_dk_travel_plan_create(
            C.byref(d), self.njob, _ptr(g["job_k"]), _ptr(g["scx"]), _ptr(g["scz"]),
            _ptr(g["rec_off"]), _ptr(g["rcx"]), _ptr(g["rcz"]), C.byref(self._h))
        _lib.check(rc)

    def run(self, pvRc, want_ttn=False, out=None, stream=None):
        """pvRc: device float64 tensor holding (nx*ny, kmax) in Fortran order
        (the stage's output; shape (kmax, nx*ny) in C order).  Enqueues and
        returns dict(times (nrec,) f32, status (njob,) i32, ttn (njob, nnx,
        nnz) f32 or None) on the device."""
        import torch
        d = self.desc
        ncol = (d.nvx + 2) * (d.nvz + 2)
        if pvRc.dtype != torch.float64 or pvRc.numel() != ncol * d.kmax or \
                not pvRc.is_contiguous():
            raise ValueError(f"pvRc must be a contiguous float64 tensor of {d.kmax} x {ncol}")
        dev = pvRc.device
        if out is None:
            nnz, nnx = grid_shape(d)
            out = dict(times=torch.empty(max(self.nrec, 1), dtype=torch.float32, device=dev),
                       status=torch.empty(max(self.njob, 1), dtype=torch.int32, device=dev),
                       ttn=torch.empty((self.njob, nnx, nnz), dtype=torch.float32, device=dev)
                       if want_ttn else None)
        rc = _lib.lib().dazi_dk_travel_run_dev(self._h, _tptr(pvRc), _tptr(out["times"]),
                                              _tptr(out.get("ttn")), _tptr(out["status"]),
                                              _stream_ptr(stream))
        _lib.check(rc)
        return dict(times=out["times"][:self.nrec], status=out["status"][:self.njob],
                    ttn=out.get("ttn"))

    def close(self):
        if self._h:
            _lib.lib().dazi_dk_travel_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def travel_times(d, job_k, scx, scz, rec_off, rcx, rcz, pvRc, want_ttn=False):
    """Host arrays in and out (dazi_dk_travel).  pvRc (nx*ny, kmax) float64 in
    Fortran order (or its (kmax, nx*ny) C-order view).  Returns times (nrec,)
    f32, status (njob,) i32 and, with want_ttn, ttn (njob, nnx, nnz) f32 (each
    job's ttn(nnz, nnx) in Fortran order)."""
    g = _geometry(job_k, scx, scz, rec_off, rcx, rcz)
    ncol = (d.nvx + 2) * (d.nvz + 2)
    pv = np.asarray(pvRc, np.float64)
    if pv.shape == (ncol, d.kmax):
        pv = pv.T
    pv = np.ascontiguousarray(pv, np.float64).reshape(-1)
    if pv.size != ncol * d.kmax:
        raise ValueError(f"pvRc holds {pv.size} values, expected {ncol * d.kmax}")
    njob = g["job_k"].size
    nrec = int(g["rec_off"][-1])
    nnz, nnx = grid_shape(d)
    times = np.empty(max(nrec, 1), np.float32)
    status = np.empty(max(njob, 1), np.int32)
    ttn = np.empty((njob, nnx, nnz), np.float32) if want_ttn else None
    rc = _lib.lib().dazi_dk_travel(C.byref(d), njob, _ptr(g["job_k"]), _ptr(g["scx"]),
                                   _ptr(g["scz"]), _ptr(g["rec_off"]), _ptr(g["rcx"]),
                                   _ptr(g["rcz"]), _ptr(pv), _ptr(times), _ptr(ttn),
                                   _ptr(status))
    _lib.check(rc)
    return times[:nrec], status[:njob], ttn
