"""Host mirror of the reference's depth-kernel stage interface.

``depthkernel`` keeps the argument list, argument meaning, in-place output
semantics and log lines of the reference subroutine
(code/src_inv_iso_joint/CalSurfG.f90:1-113); the work runs in the HIP library
through the C ABI.  ``iso_dev`` / ``layered_dev`` are the device-tensor entry
points used by the PyTorch-ROCm wrapper and by the benchmark.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib

_BACKEND_NAME = "hip"


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _fortran(a, dtype, shape, name):
    a = np.asarray(a)
    if a.dtype != dtype:
        raise TypeError(f"{name} must be {np.dtype(dtype).name}, got {a.dtype}")
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{name} has {a.size} elements, expected {shape}")
    if not (a.flags.f_contiguous or a.ndim <= 1 or a.flags.c_contiguous and a.ndim == 1):
        raise ValueError(f"{name} must be Fortran-contiguous")
    return a


def _env_int(name, default):
    import os
    try:
        v = int(os.environ.get(name, "").strip() or default)
    except ValueError:
        v = default
    return max(v, 1)


def _dump(var, label, log, arrays):
    """Writes arrays {file name: Fortran-ordered array} as raw stream files into
    the directory named by environment variable `var`, as the reference's
    depthkernel_dump_arrays (CalSurfG.f90:115-140) and depthkernelTI_dump_arrays
    (depthkernelTI.f90:195-213) do."""
    import os
    d = os.environ.get(var)
    if d is None:
        return
    d = d.strip()
    if log is not None:
        print(f"{label}{d}", file=log)
    if not d:
        return
    os.makedirs(d, exist_ok=True)
    for name, a in arrays.items():
        np.asarray(a).ravel(order="F").tofile(os.path.join(d, name))


def depthkernel(nx, ny, nz, vel, pvRc, sen_vsRc, sen_vpRc, sen_rhoRc, iwave, igr, kmaxRc, tRc,
                depz, minthk, *, vp=None, rho=None, log=sys.stdout):
    """Reference: depthkernel (CalSurfG.f90:1-113).

    vel (nx,ny,nz) float32, depz (nz,) float32, tRc (kmaxRc,) float64 ascending;
    fills pvRc (nx*ny,kmaxRc) and sen_*Rc (nx*ny,kmaxRc,nz) float64, all
    Fortran-ordered, in place.  Optional vp/rho (nx,ny,nz) replace the
    Brocher relation (the reference always uses Brocher).  Raises DaziDkError
    with the ABI code on failure (the reference stops with a code).
    """
    ncol = nx * ny
    vel = _fortran(vel, np.float32, (nx, ny, nz), "vel")
    depz = np.ascontiguousarray(depz, np.float32)
    tRc = np.ascontiguousarray(tRc, np.float64)
    pvRc = _fortran(pvRc, np.float64, (ncol, kmaxRc), "pvRc")
    outs = [sen_vsRc, sen_vpRc, sen_rhoRc]
    if any(o is not None for o in outs):
        outs = [_fortran(o, np.float64, (ncol, kmaxRc, nz), n)
                for o, n in zip(outs, ("sen_vsRc", "sen_vpRc", "sen_rhoRc"))]
    if log is not None:
        print(f"Depth kernel backend: {_BACKEND_NAME}", file=log)
        print(f"Depth kernel grid/periods: nx={nx} ny={ny} nz={nz} kmax={kmaxRc}", file=log)
    mode = 0
    if vp is not None or rho is not None:
        vp = _fortran(vp, np.float32, (nx, ny, nz), "vp")
        rho = _fortran(rho, np.float32, (nx, ny, nz), "rho")
        mode = 1
    d = _lib.Desc(nx, ny, nz, kmaxRc, iwave, igr, _env_int("DAZI_PDSURF_TILE_COLUMNS", 64), mode,
                  float(minthk), 0)
    rc = _lib.lib().dazi_dk_iso(C.byref(d), _ptr(vel), _ptr(vp), _ptr(rho), _ptr(depz),
                                _ptr(tRc), _ptr(pvRc), *[_ptr(o) for o in outs])
    _lib.check(rc)
    if outs[0] is not None:
        _dump("DAZI_DEPTHKERNEL_DUMP_DIR", "Depth kernel dump dir: ", log,
              {"pv_depthkernel.bin": pvRc, "sen_vs_depthkernel.bin": outs[0],
               "sen_vp_depthkernel.bin": outs[1], "sen_rho_depthkernel.bin": outs[2]})


def run_depthkernel(vel, tRc, depz, minthk, vp=None, rho=None, kernels=True):
    """Convenience wrapper: allocates the outputs of depthkernel and returns them
    as (pv (ncol,K), sen_vs, sen_vp, sen_rho (ncol,K,nz)) Fortran-ordered arrays."""
    vel = np.asfortranarray(vel, np.float32)
    nx, ny, nz = vel.shape
    K = len(tRc)
    pv = np.zeros((nx * ny, K), np.float64, order="F")
    sens = [np.zeros((nx * ny, K, nz), np.float64, order="F") for _ in range(3)] if kernels \
        else [None] * 3
    depthkernel(nx, ny, nz, vel, pv, *sens, 2, 0, K, tRc, depz, minthk,
                vp=None if vp is None else np.asfortranarray(vp, np.float32),
                rho=None if rho is None else np.asfortranarray(rho, np.float32), log=None)
    return (pv, *sens) if kernels else pv


# ----------------------------------------------------------------- device tensors
def _tptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def iso_dev(vel, depz, periods, minthk, vp=None, rho=None, kernels=True, tune=None, out=None,
            stream=None):
    """Node-model stage on device tensors.

    vel/vp/rho: (nz, ncol) float32 (memory order of the reference's vel(nx,ny,nz));
    depz (nz,) float32, periods (K,) float64, all on the GPU.
    Returns dict(pv (K, ncol), sen_vs/sen_vp/sen_rho (nz, K, ncol) or None, status (ncol,)).
    """
    import torch
    nz, ncol = vel.shape
    K = periods.numel()
    dev = vel.device
    if out is None:
        out = dict(pv=torch.empty((K, ncol), dtype=torch.float64, device=dev),
                   status=torch.empty(ncol, dtype=torch.int32, device=dev))
        if kernels:
            for k in ("sen_vs", "sen_vp", "sen_rho"):
                out[k] = torch.empty((nz, K, ncol), dtype=torch.float64, device=dev)
    mode = 1 if vp is not None else 0
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, mode, float(minthk), 0)
    t = tune if tune is not None else None
    rc = _lib.lib().dazi_dk_iso_dev(
        C.byref(d), None if t is None else C.byref(t), _tptr(vel), _tptr(vp), _tptr(rho),
        _tptr(depz), _tptr(periods), _tptr(out["pv"]),
        _tptr(out.get("sen_vs")) if kernels else None,
        _tptr(out.get("sen_vp")) if kernels else None,
        _tptr(out.get("sen_rho")) if kernels else None, _tptr(out["status"]),
        _stream_ptr(stream))
    _lib.check(rc)
    return out


class LayeredBatch:
    """Ragged layered batch resident on the GPU (the batched solver's input)."""

    def __init__(self, nl, nk, thk, vp, vs, rho, periods, device="cuda"):
        """Model arrays may be float32 (the reference's precision) or float64
        (flattened in float64: the differentiable path)."""
        import torch
        nl = np.asarray(nl, np.int32)
        nk = np.asarray(nk, np.int32)
        self.ncol = int(nl.size)
        self.max_layers = int(nl.max()) if nl.size else 1
        self.nsolves = int(nk.sum())
        loff = np.concatenate([[0], np.cumsum(nl, dtype=np.int64)[:-1]]).astype(np.int64)
        toff = np.concatenate([[0], np.cumsum(nk, dtype=np.int64)[:-1]]).astype(np.int64)
        solve_col = np.repeat(np.arange(self.ncol, dtype=np.int32), nk)
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)  # noqa
        self.nl = T(nl, torch.int32)
        self.nk = T(nk, torch.int32)
        self.loff = T(loff, torch.int64)
        self.toff = T(toff, torch.int64)
        self.solve_col = T(solve_col, torch.int32)
        self.f64 = np.asarray(vs).dtype == np.float64
        mt = torch.float64 if self.f64 else torch.float32
        self.thk = T(thk, mt)
        self.vp = T(vp, mt)
        self.vs = T(vs, mt)
        self.rho = T(rho, mt)
        self.periods = T(periods, torch.float64)
        self.nl_host = nl
        self.nk_host = nk
        self.loff_host = loff
        self.toff_host = toff
        self.device = device

    def alloc_outputs(self, kernels=False):
        import torch
        out = dict(c=torch.empty(self.nsolves, dtype=torch.float64, device=self.device),
                   c_unrounded=torch.empty(self.nsolves, dtype=torch.float64,
                                           device=self.device),
                   status=torch.empty(self.ncol, dtype=torch.int32, device=self.device))
        if kernels:
            for k in ("dc_dvs", "dc_dvp", "dc_drho"):
                out[k] = torch.zeros((self.max_layers, self.nsolves), dtype=torch.float64,
                                     device=self.device)
        return out


def layered_dev(batch: LayeredBatch, kernels=False, tune=None, out=None, stream=None,
                vs=None, vp=None, rho=None, thk=None):
    """Root search (+ refined-layer kernels) of a ragged layered batch.

    Returns dict(c (nsolves,) float32-rounded, c_unrounded (nsolves,), status (ncol,),
    dc_dvs/dc_dvp/dc_drho (max_layers, nsolves)).
    vs/vp/rho/thk override the batch's model tensors (same layout) when given.
    """
    if out is None:
        out = batch.alloc_outputs(kernels)
    b = _layered_struct(batch, vs, vp, rho, thk)
    b.c = out["c"].data_ptr()
    b.status = out["status"].data_ptr()
    if kernels:
        b.dc_dvs = out["dc_dvs"].data_ptr()
        b.dc_dvp = out["dc_dvp"].data_ptr()
        b.dc_drho = out["dc_drho"].data_ptr()
    if "c_unrounded" in out:
        b.c_unrounded = out["c_unrounded"].data_ptr()
    rc = _lib.lib().dazi_dk_layered_dev(C.byref(b), None if tune is None else C.byref(tune),
                                        _stream_ptr(stream))
    _lib.check(rc)
    return out


def _layered_struct(batch, vs=None, vp=None, rho=None, thk=None):
    vs = batch.vs if vs is None else vs
    vp = batch.vp if vp is None else vp
    rho = batch.rho if rho is None else rho
    thk = batch.thk if thk is None else thk
    for t in (vs, vp, rho, thk):
        if t.dtype != batch.thk.dtype or not t.is_contiguous() or t.numel() != batch.thk.numel():
            raise ValueError("model tensors must match the batch's dtype, size and layout")
    b = _lib.Layered()
    b.ncol, b.max_layers, b.nsolves = batch.ncol, batch.max_layers, batch.nsolves
    b.nl, b.loff = batch.nl.data_ptr(), batch.loff.data_ptr()
    b.nk, b.toff = batch.nk.data_ptr(), batch.toff.data_ptr()
    b.solve_col = batch.solve_col.data_ptr()
    b.periods = batch.periods.data_ptr()
    b.nlayers_total = int(batch.nl_host.sum())
    if batch.f64:
        b.thk64, b.vp64, b.vs64, b.rho64 = (thk.data_ptr(), vp.data_ptr(), vs.data_ptr(),
                                            rho.data_ptr())
    else:
        b.thk, b.vp, b.vs, b.rho = (thk.data_ptr(), vp.data_ptr(), vs.data_ptr(), rho.data_ptr())
    return b


def layered_vjp(batch: LayeredBatch, c_unrounded, grad_c, vs=None, vp=None, rho=None,
                thk=None, stream=None):
    """grad_x[layer] = sum over the column's periods of grad_c * dc/dx, x in (vs, vp, rho,
    thk), by implicit differentiation at the roots c_unrounded of the same model.
    Returns four float64 tensors laid out like the model (thickness gradient 0 for
    the half-space)."""
    import torch
    dev = batch.thk.device
    g = [torch.empty(batch.thk.numel(), dtype=torch.float64, device=dev) for _ in range(4)]
    b = _layered_struct(batch, vs, vp, rho, thk)
    b.c_unrounded = c_unrounded.data_ptr()
    gc = grad_c.to(torch.float64).contiguous()
    rc = _lib.lib().dazi_dk_layered_vjp_dev(C.byref(b), C.c_void_p(gc.data_ptr()),
                                            C.c_void_p(g[0].data_ptr()),
                                            C.c_void_p(g[1].data_ptr()),
                                            C.c_void_p(g[2].data_ptr()),
                                            C.c_void_p(g[3].data_ptr()), _stream_ptr(stream))
    _lib.check(rc)
    return g[0], g[1], g[2], g[3]


def depthkernelTI(nx, ny, nz, vel, pvRc, iwave, igr, kmaxRc, tRc, depz, minthk, Lsen_Gsc, *,
                  log=sys.stdout):
    """Reference: depthkernelTI (depthkernelTI.f90:2-193).

    vel (nx,ny,nz) float32; fills pvRc (nx*ny,kmaxRc) float64 and Lsen_Gsc
    (nx*ny,kmaxRc,nz-1) float32 (zeroed first, as the reference does), in place.
    """
    ncol = nx * ny
    vel = _fortran(vel, np.float32, (nx, ny, nz), "vel")
    depz = np.ascontiguousarray(depz, np.float32)
    tRc = np.ascontiguousarray(tRc, np.float64)
    pvRc = _fortran(pvRc, np.float64, (ncol, kmaxRc), "pvRc")
    Lsen_Gsc = _fortran(Lsen_Gsc, np.float32, (ncol, kmaxRc, nz - 1), "Lsen_Gsc")
    if log is not None:
        print(f"Depth kernel TI backend: {_BACKEND_NAME}", file=log)
        print(f"Depth kernel TI grid/periods: nx={nx} ny={ny} nz={nz} kmax={kmaxRc}", file=log)
    d = _lib.Desc(nx, ny, nz, kmaxRc, iwave, igr,
                  _env_int("DAZI_DEPTHKERNEL_TI_TILE_COLUMNS", 64), 0, float(minthk), 0)
    rc = _lib.lib().dazi_dk_ti(C.byref(d), _ptr(vel), _ptr(depz), _ptr(tRc), _ptr(pvRc),
                               _ptr(Lsen_Gsc))
    _lib.check(rc)
    _dump("DAZI_DEPTHKERNEL_TI_DUMP_DIR", "Depth kernel TI dump dir: ", log,
          {"pv_ti.bin": pvRc, "lsen_gsc_ti.bin": Lsen_Gsc})


def run_depthkernelTI(vel, tRc, depz, minthk):
    """Allocating wrapper: returns (pv (ncol,K), Lsen (ncol,K,nz-1)) Fortran-ordered."""
    vel = np.asfortranarray(vel, np.float32)
    nx, ny, nz = vel.shape
    K = len(tRc)
    pv = np.zeros((nx * ny, K), np.float64, order="F")
    ls = np.zeros((nx * ny, K, nz - 1), np.float32, order="F")
    depthkernelTI(nx, ny, nz, vel, pv, 2, 0, K, tRc, depz, minthk, ls, log=None)
    return pv, ls


def stage_dev(vel, depz, periods, minthk, vp=None, rho=None, iso=True, ti=True, tune=None,
              out=None, stream=None):
    """Both subroutines' outputs on device tensors from one root search.

    vel/vp/rho (nz, ncol) float32; returns dict(pv (K,ncol), sen_vs/sen_vp/sen_rho
    (nz,K,ncol) when iso, lsen (nz-1,K,ncol) float32 when ti, status (ncol,)).
    """
    import torch
    nz, ncol = vel.shape
    K = periods.numel()
    dev = vel.device
    if out is None:
        out = dict(pv=torch.empty((K, ncol), dtype=torch.float64, device=dev),
                   status=torch.empty(ncol, dtype=torch.int32, device=dev))
        if iso:
            for k in ("sen_vs", "sen_vp", "sen_rho"):
                out[k] = torch.empty((nz, K, ncol), dtype=torch.float64, device=dev)
        if ti:
            out["lsen"] = torch.empty((nz - 1, K, ncol), dtype=torch.float32, device=dev)
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, 1 if vp is not None else 0, float(minthk), 0)
    rc = _lib.lib().dazi_dk_stage_dev(
        C.byref(d), None if tune is None else C.byref(tune), _tptr(vel), _tptr(vp), _tptr(rho),
        _tptr(depz), _tptr(periods), _tptr(out["pv"]),
        _tptr(out.get("sen_vs")) if iso else None, _tptr(out.get("sen_vp")) if iso else None,
        _tptr(out.get("sen_rho")) if iso else None, _tptr(out.get("lsen")) if ti else None,
        _tptr(out["status"]), _stream_ptr(stream))
    _lib.check(rc)
    return out


def effective_dev(vel, sen_vs, sen_vp, sen_rho, grad_c=None, eff=True, grad=True, out=None,
                  stream=None):
    """Effective dc/dVs of the G builder's Vs row under the Brocher relation
    (code/src_inv_iso_joint/CalSurfGAniso_Joint.f90:750-765) and its contraction
    with grad_c over the periods (the VJP with respect to node Vs).

    vel (nz, ncol) f32, sen_* (nz, K, ncol) f64, grad_c (K, ncol) f64; returns
    dict(eff (nz-1, K, ncol), grad_vs (nz, ncol)) as requested."""
    import torch
    nz, K, ncol = sen_vs.shape
    dev = sen_vs.device
    if out is None:
        out = {}
        if eff:
            out["eff"] = torch.empty((nz - 1, K, ncol), dtype=torch.float64, device=dev)
        if grad and grad_c is not None:
            out["grad_vs"] = torch.empty((nz, ncol), dtype=torch.float64, device=dev)
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, 0, 1.0, 0)
    rc = _lib.lib().dazi_dk_effective_dev(
        C.byref(d), _tptr(vel), _tptr(sen_vs), _tptr(sen_vp), _tptr(sen_rho), _tptr(grad_c),
        _tptr(out.get("eff")), _tptr(out.get("grad_vs")), _stream_ptr(stream))
    _lib.check(rc)
    return out
