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


def _env_on(name):
    import os
    v = os.environ.get(name, "").strip()
    return bool(v) and v != "0"


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


def _check_dev(t, dtype, shape, name, optional=False):
    """Raises ValueError unless t is a contiguous GPU tensor of the given dtype
    and shape (the C ABI reads raw pointers: a wrong dtype, a strided view or a
    short buffer would read or write out of bounds)."""
    import torch
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the GPU")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on the GPU, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _check_node_inputs(vel, depz, periods, vp, rho):
    """Node-model inputs of the device entry points: vel (nz, ncol) and
    optional vp, rho (both or neither) float32, depz (nz,) float32, periods
    (K,) float64; all contiguous on one GPU."""
    import torch
    if not isinstance(vel, torch.Tensor) or vel.dim() != 2:
        raise ValueError("vel must be a (nz, ncol) float32 GPU tensor")
    nz, ncol = vel.shape
    _check_dev(vel, torch.float32, (nz, ncol), "vel")
    _check_dev(depz, torch.float32, (nz,), "depz")
    if not isinstance(periods, torch.Tensor) or periods.dim() != 1:
        raise ValueError("periods must be a (K,) float64 GPU tensor")
    _check_dev(periods, torch.float64, (periods.numel(),), "periods")
    if (vp is None) != (rho is None):
        raise ValueError("vp and rho must be given together")
    _check_dev(vp, torch.float32, (nz, ncol), "vp", optional=True)
    _check_dev(rho, torch.float32, (nz, ncol), "rho", optional=True)
    for t in (depz, periods, vp, rho):
        if t is not None and t.device != vel.device:
            raise ValueError("all inputs must be on the same GPU")
    return nz, ncol, periods.numel()


# The depz tensor of the last device call on each GPU, with its address and
# torch version counter at that call.
_DEPZ_SEEN = {}


def _depz_flags(depz):
    """DAZI_DK_DEPZ_UNCHANGED when depz is the very tensor of the previous call
    on its GPU and torch has recorded no write to it since; the library then
    skips reading depz back to check its cached node geometry.  Holding the
    tensor keeps its memory from being handed to another tensor meanwhile."""
    key = (depz.data_ptr(), depz._version)
    dev = depz.device.index
    prev = _DEPZ_SEEN.get(dev)
    _DEPZ_SEEN[dev] = (depz, key)
    same = prev is not None and prev[0] is depz and prev[1] == key
    return _lib.DEPZ_UNCHANGED if same else 0


def _check_out(out, spec):
    """Caller-given outputs: spec {key: (dtype, shape)}; every key must be present."""
    for k, (dt, shape) in spec.items():
        if k not in out:
            raise ValueError(f"out is missing {k!r}")
        _check_dev(out[k], dt, shape, f"out[{k!r}]")


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
    nz, ncol, K = _check_node_inputs(vel, depz, periods, vp, rho)
    dev = vel.device
    spec = dict(pv=(torch.float64, (K, ncol)), status=(torch.int32, (ncol,)))
    if kernels:
        spec.update({k: (torch.float64, (nz, K, ncol)) for k in ("sen_vs", "sen_vp", "sen_rho")})
    if out is None:
        out = {k: torch.empty(shape, dtype=dt, device=dev) for k, (dt, shape) in spec.items()}
    _check_out(out, spec)
    mode = 1 if vp is not None else 0
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, mode, float(minthk), _depz_flags(depz))
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
    # NaN at periods without a phase velocity, as the reference leaves there
    # (tregn96 at c = 0), on request; 0 otherwise
    flags = _lib.TI_FAILED_NAN if _env_on("DAZI_DEPTHKERNEL_TI_FAILED_NAN") else 0
    d = _lib.Desc(nx, ny, nz, kmaxRc, iwave, igr,
                  _env_int("DAZI_DEPTHKERNEL_TI_TILE_COLUMNS", 64), 0, float(minthk), flags)
    L = _lib.lib()
    if log is not None:
        L.dazi_dk_enable_timing(1)
    try:
        rc = L.dazi_dk_ti(C.byref(d), _ptr(vel), _ptr(depz), _ptr(tRc), _ptr(pvRc),
                          _ptr(Lsen_Gsc))
        tm = _lib.last_timings() if (log is not None and rc == 0) else []
    finally:
        if log is not None:
            L.dazi_dk_enable_timing(0)
    _lib.check(rc)
    if len(tm) >= 7:
        # the reference's internal timing lines (depthkernelTI.f90:170-186)
        lref = int(round(tm[5]))
        print("  depthkernelTI internal timing:", file=log)
        print(f"    refine+surfdisp96 phase velocity = {(tm[0] + tm[1]) * 1e-3:12.3f} s", file=log)
        print(f"    tregn96 sensitivity kernel = {tm[3] * 1e-3:12.3f} s", file=log)
        print(f"    splice refined kernel = {0.0:12.3f} s", file=log)
        print(f"    refined sublayers rmax min/max/avg = {lref}/{lref}/{float(lref):10.2f}",
              file=log)
        print(f"    observed worker threads = {int(round(tm[6]))}", file=log)
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
    nz, ncol, K = _check_node_inputs(vel, depz, periods, vp, rho)
    dev = vel.device
    spec = dict(pv=(torch.float64, (K, ncol)), status=(torch.int32, (ncol,)))
    if iso:
        spec.update({k: (torch.float64, (nz, K, ncol)) for k in ("sen_vs", "sen_vp", "sen_rho")})
    if ti:
        spec["lsen"] = (torch.float32, (nz - 1, K, ncol))
    if out is None:
        out = {k: torch.empty(shape, dtype=dt, device=dev) for k, (dt, shape) in spec.items()}
    _check_out(out, spec)
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, 1 if vp is not None else 0, float(minthk),
                  _depz_flags(depz))
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
    if not isinstance(sen_vs, torch.Tensor) or sen_vs.dim() != 3:
        raise ValueError("sen_vs must be a (nz, K, ncol) float64 GPU tensor")
    nz, K, ncol = sen_vs.shape
    dev = sen_vs.device
    _check_dev(vel, torch.float32, (nz, ncol), "vel")
    for t, n in ((sen_vs, "sen_vs"), (sen_vp, "sen_vp"), (sen_rho, "sen_rho")):
        _check_dev(t, torch.float64, (nz, K, ncol), n)
    _check_dev(grad_c, torch.float64, (K, ncol), "grad_c", optional=True)
    shapes = dict(eff=(torch.float64, (nz - 1, K, ncol)), grad_vs=(torch.float64, (nz, ncol)))
    if out is None:
        keys = [k for k, want in (("eff", eff), ("grad_vs", grad and grad_c is not None)) if want]
        out = {k: torch.empty(shapes[k][1], dtype=torch.float64, device=dev) for k in keys}
    # with caller-given outputs, the keys present name the outputs computed
    if "grad_vs" in out and grad_c is None:
        raise ValueError("out['grad_vs'] needs grad_c")
    _check_out(out, {k: v for k, v in shapes.items() if k in out})
    for t in (vel, sen_vp, sen_rho, grad_c, *out.values()):
        if t is not None and t.device != dev:
            raise ValueError("all tensors must be on the same GPU")
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, 0, 1.0, 0)
    rc = _lib.lib().dazi_dk_effective_dev(
        C.byref(d), _tptr(vel), _tptr(sen_vs), _tptr(sen_vp), _tptr(sen_rho), _tptr(grad_c),
        _tptr(out.get("eff")), _tptr(out.get("grad_vs")), _stream_ptr(stream))
    _lib.check(rc)
    return out


def node_layers(depz, minthk):
    """Refined layers per column (half-space included) of the node geometry."""
    d = np.ascontiguousarray(np.asarray(depz.cpu() if hasattr(depz, "cpu") else depz,
                                        np.float32))
    n = C.c_int32(0)
    _lib.check(_lib.lib().dazi_dk_node_layers(d.size, _ptr(d), float(minthk), C.byref(n)))
    return n.value


def ti_kernels_dev(vel, depz, periods, minthk, vp=None, rho=None, lsen=True, tune=None,
                   out=None, stream=None):
    """The TI stage with the per-refined-layer kernels depthkernelTI splices
    (depthkernelTI.f90:152-162) on device tensors.

    vel/vp/rho (nz, ncol) float32, depz (nz,) float32, periods (K,) float64.
    Returns dict(pv (K, ncol) f64, dc_dA, dc_dL (L, K, ncol) f32, lsen (nz-1, K,
    ncol) f32 when lsen, status (ncol,) i32), L = node_layers(depz, minthk)."""
    import torch
    nz, ncol, K = _check_node_inputs(vel, depz, periods, vp, rho)
    L = node_layers(depz, minthk)
    dev = vel.device
    spec = dict(pv=(torch.float64, (K, ncol)), status=(torch.int32, (ncol,)),
                dc_dA=(torch.float32, (L, K, ncol)), dc_dL=(torch.float32, (L, K, ncol)))
    if lsen:
        spec["lsen"] = (torch.float32, (nz - 1, K, ncol))
    if out is None:
        out = {k: torch.empty(shape, dtype=dt, device=dev) for k, (dt, shape) in spec.items()}
    _check_out(out, spec)
    d = _lib.Desc(ncol, 1, nz, K, 2, 0, 64, 1 if vp is not None else 0, float(minthk),
                  _depz_flags(depz))
    rc = _lib.lib().dazi_dk_ti_kernels_dev(
        C.byref(d), None if tune is None else C.byref(tune), _tptr(vel), _tptr(vp), _tptr(rho),
        _tptr(depz), _tptr(periods), _tptr(out["pv"]), _tptr(out.get("lsen")) if lsen else None,
        _tptr(out["dc_dA"]), _tptr(out["dc_dL"]), _tptr(out["status"]), _stream_ptr(stream))
    _lib.check(rc)
    return out


def ti_vjp_dev(dc_dA, dc_dL, grad_c, stream=None):
    """grad_A, grad_L (L, ncol) float64: sum over periods of grad_c * dc/dA, dc/dL
    per refined layer (dc_dA, dc_dL (L, K, ncol) float32; grad_c (K, ncol))."""
    import torch
    if not isinstance(dc_dA, torch.Tensor) or dc_dA.dim() != 3:
        raise ValueError("dc_dA must be an (L, K, ncol) float32 GPU tensor")
    L, K, ncol = dc_dA.shape
    _check_dev(dc_dA, torch.float32, (L, K, ncol), "dc_dA")
    _check_dev(dc_dL, torch.float32, (L, K, ncol), "dc_dL")
    gc = grad_c.to(torch.float64).contiguous()
    _check_dev(gc, torch.float64, (K, ncol), "grad_c")
    gA = torch.empty((L, ncol), dtype=torch.float64, device=dc_dA.device)
    gL = torch.empty_like(gA)
    rc = _lib.lib().dazi_dk_ti_vjp_dev(ncol, L, K, _tptr(dc_dA), _tptr(dc_dL), _tptr(gc),
                                       _tptr(gA), _tptr(gL), _stream_ptr(stream))
    _lib.check(rc)
    return gA, gL
