"""ORACLE — test infrastructure only (tests/, smoke() and bench.py's cpu_baseline leg).

ctypes access to two CPU implementations of the reference's per-column stage
and of the traveltime stage that consumes its phase velocities:

* ``liborc_port.so`` — this repository's C restatements (``surfdisp96_port.c``,
  ``fmm_port.c``);
* ``_ref/libdazi_ref.so`` — the reference's own Fortran (surfdisp96.f, tregn96.f,
  DepthKernel1D.f90, CalSurfG.f90, depthkernel_ti_worker.f90,
  FwdAzimuthalAniMap.f90) compiled in place with flang by ``oracle/Makefile``,
  with this directory's ``fmm_ref_driver.f90``.  Present only where it was built.

Array conventions follow the reference: Fortran column-major, float32 models,
float64 periods and outputs.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PORT = os.path.join(_HERE, "liborc_port.so")
_REF = os.path.join(_HERE, "_ref", "libdazi_ref.so")

_port = None
_ref = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

NL = 200
NP = 160


def port():
    global _port
    if _port is None:
        if not os.path.exists(_PORT):
            raise RuntimeError(f"oracle port library missing: {_PORT} (run make -C oracle)")
        L = C.CDLL(_PORT)
        L.orc_gtsolh.restype = C.c_float
        L.orc_gtsolh.argtypes = [C.c_float, C.c_float]
        L.orc_dltar4.restype = C.c_double
        L.orc_dltar4.argtypes = [C.c_double, C.c_double, f32p, f32p, f32p, f32p, C.c_int, C.c_int]
        L.orc_surfdisp96_x.restype = C.c_int
        L.orc_surfdisp96_x.argtypes = [f32p, f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, f64p,
                                       f64p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_brocher.argtypes = [C.c_int, f32p, f32p, f32p]
        L.orc_refine.restype = C.c_int
        L.orc_refine.argtypes = [C.c_float, C.c_int, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                 f32p, C.c_void_p]
        L.orc_column_fd.restype = C.c_int
        L.orc_column_fd.argtypes = [C.c_int, f32p, C.c_void_p, C.c_void_p, C.c_int, f64p, f32p,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_double,
                                    C.c_int, C.c_int, f64p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_depthkernel.argtypes = [C.c_int, C.c_int, C.c_int, f32p, C.c_void_p, C.c_void_p,
                                      f64p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, f64p,
                                      f32p, C.c_float, C.c_int]
        L.orc_forward_layered.argtypes = [C.c_int, i32p, i64p, i32p, i64p, f32p, f32p, f32p,
                                          f32p, f64p, f64p, C.c_void_p, C.c_int]
        L.orc_forward_layered_x.argtypes = [C.c_int, i32p, i64p, i32p, i64p, f32p, f32p, f32p,
                                            f32p, f64p, f64p, C.c_void_p, C.c_int, C.c_double,
                                            C.c_int]
        L.orc_max_threads.restype = C.c_int
        _port = L
    return _port


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference library missing: {_REF} (run make -C oracle ref)")
        _ref = C.CDLL(_REF)
        _ref.dltar4_.restype = C.c_double
    return _ref


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size < n:
        b = np.zeros(n, np.float32)
        b[: a.size] = a
        return b
    return a


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size < n:
        b = np.zeros(n, np.float64)
        b[: a.size] = a
        return b
    return a


# ----------------------------------------------------------------- models
def brocher(vs, lib="port"):
    vs = _f32(vs)
    n = vs.size
    vp = np.zeros(n, np.float32)
    rho = np.zeros(n, np.float32)
    if lib == "port":
        port().orc_brocher(n, vs, vp, rho)
    else:
        ref().depthkernel_empirical_vp_rho_(C.byref(C.c_int(n)), _ptr(vs), _ptr(vp), _ptr(rho))
    return vp, rho


def refine(minthk, depz, vp, vs, rho, lib="port"):
    """Node model -> refined layered model (thk, vp, vs, rho), half-space last."""
    depz, vp, vs, rho = (_f32(x) for x in (depz, vp, vs, rho))
    nz = depz.size
    out = [np.zeros(NL, np.float32) for _ in range(5)]
    if lib == "port":
        L = port().orc_refine(C.c_float(minthk), nz, depz, vp, vs, rho, out[1], out[2], out[3],
                              out[0], None)
        if L < 0:
            raise ValueError("refined model exceeds NL")
    else:
        Lc = C.c_int(0)
        ref().refinegrid2layermdl_(C.byref(C.c_float(minthk)), C.byref(C.c_int(nz)), _ptr(depz),
                                   _ptr(vp), _ptr(vs), _ptr(rho), C.byref(Lc), _ptr(out[4]),
                                   _ptr(out[1]), _ptr(out[2]), _ptr(out[3]), _ptr(out[0]))
        L = Lc.value
    return out[0][:L].copy(), out[1][:L].copy(), out[2][:L].copy(), out[3][:L].copy()


# ----------------------------------------------------------------- forward
def surfdisp96(thk, vp, vs, rho, periods, iflsph=1, lib="port", tol_rel=1e-6, round_f32=True,
               want_flat=False, want_nevals=False):
    """Rayleigh fundamental-mode phase velocities of one layered model.

    Returns cg (kmax,) float64 with zeros from the first failed period on.
    lib='ref' calls the reference's own surfdisp96 (tol_rel/round_f32 fixed).
    """
    n = len(thk)
    periods = _f64(periods)
    kmax = periods.size
    if lib == "port":
        cg = np.zeros(kmax, np.float64)
        flat = np.zeros(4 * n, np.float32) if want_flat else None
        nev = C.c_int(0)
        port().orc_surfdisp96_x(_f32(thk), _f32(vp), _f32(vs), _f32(rho), n, iflsph, kmax,
                                periods, cg, tol_rel, 1 if round_f32 else 0,
                                _ptr(flat) if want_flat else None, C.byref(nev))
        out = [cg]
        if want_flat:
            out.append(flat.reshape(4, n))
        if want_nevals:
            out.append(nev.value)
        return out[0] if len(out) == 1 else tuple(out)
    cg = np.zeros(NP, np.float64)
    ref().surfdisp96_(_ptr(_f32(thk, NL)), _ptr(_f32(vp, NL)), _ptr(_f32(vs, NL)),
                      _ptr(_f32(rho, NL)), C.byref(C.c_int(n)), C.byref(C.c_int(iflsph)),
                      C.byref(C.c_int(2)), C.byref(C.c_int(1)), C.byref(C.c_int(0)),
                      C.byref(C.c_int(kmax)), _ptr(_f64(periods, NP)), _ptr(cg))
    return cg[:kmax].copy()


def gtsolh(a, b, lib="port"):
    if lib == "port":
        return np.float32(port().orc_gtsolh(a, b))
    c = C.c_float(0)
    ref().gtsolh_(C.byref(C.c_float(a)), C.byref(C.c_float(b)), C.byref(c))
    return np.float32(c.value)


def dltar4(c, period, d, a, b, rho, lib="port", llw=1):
    """Secular function at phase velocity c and period, on an already-flattened model."""
    omega = 2.0 * 3.141592653589793 / period
    wvno = omega / c
    n = len(d)
    if lib == "port":
        return port().orc_dltar4(wvno, omega, _f32(d), _f32(a), _f32(b), _f32(rho), n, llw)
    z = np.zeros(NL, np.float32)
    return ref().dltar4_(C.byref(C.c_double(wvno)), C.byref(C.c_double(omega)), _ptr(_f32(d, NL)),
                         _ptr(_f32(a, NL)), _ptr(_f32(b, NL)), _ptr(_f32(rho, NL)), _ptr(z),
                         _ptr(z), _ptr(z), C.byref(C.c_int(n)), C.byref(C.c_int(llw)),
                         C.byref(C.c_double(2 * np.pi)))


# ----------------------------------------------------------------- FD depth kernels
def column_fd(vs, periods, depz, minthk, vp=None, rho=None, lib="port", dln=0.01,
              tol_rel=1e-6, round_f32=True, kernels=True, exact_step=False):
    """One column of the reference FD depth kernel.

    Returns pv (kmax,), sen_vs/sen_vp/sen_rho (kmax, nz) (None when kernels=False).
    tol_rel=1e-13, round_f32=False, exact_step=True give a tight FD oracle.
    """
    vs = _f32(vs)
    nz = vs.size
    periods = _f64(periods)
    kmax = periods.size
    depz = _f32(depz)
    pv = np.zeros(kmax, np.float64)
    if lib == "port":
        sens = [np.zeros(kmax * nz, np.float64) for _ in range(3)] if kernels else [None] * 3
        vpa = _f32(vp) if vp is not None else None
        rhoa = _f32(rho) if rho is not None else None
        port().orc_column_fd(nz, vs, _ptr(vpa) if vpa is not None else None,
                             _ptr(rhoa) if rhoa is not None else None, kmax, periods, depz,
                             C.c_float(minthk), dln, dln, dln, tol_rel, 1 if round_f32 else 0,
                             1 if exact_step else 0, pv,
                             *[_ptr(s) if s is not None else None for s in sens])
        if not kernels:
            return pv, None, None, None
        return (pv,) + tuple(s.reshape(nz, kmax).T.copy() for s in sens)
    if vp is not None:
        raise ValueError("the reference column driver derives vp/rho from vs")
    sens = [np.zeros(kmax * nz, np.float64) for _ in range(3)]
    d = C.c_float(dln)
    ref().depthkernel_column_fd_(C.byref(C.c_int(nz)), _ptr(vs), C.byref(C.c_int(kmax)),
                                 _ptr(periods), C.byref(C.c_int(2)), C.byref(C.c_int(0)),
                                 _ptr(depz), C.byref(C.c_float(minthk)), C.byref(d), C.byref(d),
                                 C.byref(d), _ptr(pv), *[_ptr(s) for s in sens])
    return (pv,) + tuple(s.reshape(nz, kmax).T.copy() for s in sens)


def depthkernel(vel, tRc, depz, minthk, lib="port", nthreads=0, kernels=True, vp=None, rho=None):
    """The stage call: vel (nx, ny, nz) -> pv (nx*ny, kmax), sen_* (nx*ny, kmax, nz).

    Arrays use Fortran order semantics (column index fastest); numpy arrays are
    returned with shape (ncol, kmax[, nz]) in Fortran memory order.
    """
    vel = np.asarray(vel, np.float32)
    nx, ny, nz = vel.shape
    velF = np.ascontiguousarray(vel.reshape(-1, order="F"))
    tRc = _f64(tRc)
    kmax = tRc.size
    ncol = nx * ny
    pv = np.zeros(ncol * kmax, np.float64)
    sens = [np.zeros(ncol * kmax * nz, np.float64) for _ in range(3)] if kernels else None
    if lib == "port":
        vpF = np.ascontiguousarray(np.asarray(vp, np.float32).reshape(-1, order="F")) \
            if vp is not None else None
        rhoF = np.ascontiguousarray(np.asarray(rho, np.float32).reshape(-1, order="F")) \
            if rho is not None else None
        port().orc_depthkernel(nx, ny, nz, velF, _ptr(vpF) if vpF is not None else None,
                               _ptr(rhoF) if rhoF is not None else None, pv,
                               *([_ptr(s) for s in sens] if kernels else [None] * 3), kmax, tRc,
                               _f32(depz), C.c_float(minthk), nthreads)
    else:
        if not kernels or vp is not None:
            raise ValueError("reference depthkernel always computes kernels from Brocher vp/rho")
        if nthreads > 0:
            os.environ["OMP_NUM_THREADS"] = str(nthreads)
        os.environ["DAZI_DEPTHKERNEL_BACKEND"] = "default_fortran_fd"
        os.environ.pop("DAZI_DEPTHKERNEL_DUMP_DIR", None)
        i = lambda v: C.byref(C.c_int(v))  # noqa: E731
        ref().depthkernel_(i(nx), i(ny), i(nz), _ptr(velF), _ptr(pv), *[_ptr(s) for s in sens],
                           i(2), i(0), i(kmax), _ptr(tRc), _ptr(_f32(depz)),
                           C.byref(C.c_float(minthk)))
    pv = pv.reshape(ncol, kmax, order="F")
    if not kernels:
        return pv
    return (pv,) + tuple(s.reshape(ncol, kmax, nz, order="F") for s in sens)


def forward_layered(nl, kk, thk, vp, vs, rho, periods, nthreads=0, tol_rel=1e-6,
                    round_f32=True):
    """Ragged batch forward: per column surfdisp96 over flat concatenated arrays.
    tol_rel=1e-13, round_f32=False give the tight roots of a finite-difference
    oracle."""
    nl = np.ascontiguousarray(nl, np.int32)
    kk = np.ascontiguousarray(kk, np.int32)
    loff = np.concatenate([[0], np.cumsum(nl)[:-1]]).astype(np.int64)
    toff = np.concatenate([[0], np.cumsum(kk)[:-1]]).astype(np.int64)
    cg = np.zeros(int(kk.sum()), np.float64)
    fail = np.zeros(nl.size, np.int32)
    port().orc_forward_layered_x(nl.size, nl, loff, kk, toff, _f32(thk), _f32(vp), _f32(vs),
                                 _f32(rho), _f64(periods), cg, _ptr(fail), nthreads, tol_rel,
                                 1 if round_f32 else 0)
    return cg, fail


def layer_kernels_fd(thk, vp, vs, rho, periods, rel=(0.005, 0.0025), nthreads=0):
    """Tight finite-difference layer kernels of ONE layered model: dc/dVs,
    dc/dVp, dc/drho (each (kmax, L)) by central differences at float32 model
    steps of rel[0] and rel[1] of each value, combined by Richardson
    extrapolation (O(h^4) truncation), with roots to 1e-13 relative and no
    float32 rounding of c.  All 12 L perturbed models run as one OpenMP batch.
    Returns (c (kmax,), [dvs, dvp, drho])."""
    base = [np.asarray(a, np.float32).copy() for a in (thk, vp, vs, rho)]
    L = base[0].size
    periods = np.asarray(periods, np.float64)
    K = periods.size
    models, steps = [base], []
    for q in (2, 1, 3):                      # vs, vp, rho
        for m in range(L):
            for r in rel:
                v0 = base[q][m]
                h = np.float32(r) * v0
                lo, hi = np.float32(v0 - h), np.float32(v0 + h)
                for v in (lo, hi):
                    mm = [a.copy() for a in base]
                    mm[q][m] = v
                    models.append(mm)
                steps.append(float(hi) - float(lo))
    n = len(models)
    cat = [np.concatenate([mm[i] for mm in models]) for i in range(4)]
    c, _ = forward_layered(np.full(n, L, np.int32), np.full(n, K, np.int32), *cat,
                           np.tile(periods, n), nthreads=nthreads, tol_rel=1e-13,
                           round_f32=False)
    c = c.reshape(n, K)
    out, i, s = [], 1, 0
    for q in range(3):
        g = np.zeros((K, L))
        for m in range(L):
            d = []
            for _ in rel:
                d.append((c[i + 1] - c[i]) / steps[s])
                i += 2
                s += 1
            g[:, m] = (4 * d[1] - d[0]) / 3
        out.append(g)
    return c[0], out


FORWARD_BATCH = os.path.join(_HERE, "_ref", "surfdisp96_forward_batch")


def forward_batch_available() -> bool:
    return os.path.exists(FORWARD_BATCH)


def write_layered_fd_input(path, thk, vp, vs, rho, periods, dln=0.01):
    """Input file of the reference's layered forward driver
    (depthkernel_lab/surfdisp96_forward_batch.f90:26-54) holding the model and
    the 6 L perturbed models of its finite-difference kernels, ordered as
    depthkernel_column_fd perturbs (DepthKernel1D.f90:146-183): for each layer i,
    vs_i -/+ 0.5 dln vs_i, then vp, then rho, in float32.  Values are written
    with 9 significant digits, so the driver reads back the same float32 values.
    Returns the number of models (1 + 6 L)."""
    f = np.float32
    base = [np.asarray(a, np.float32).copy() for a in (thk, vp, vs, rho)]
    L = base[0].size
    periods = np.asarray(periods, np.float64)
    models = [base]
    half = f(0.5) * f(dln)
    for i in range(L):
        for q in (2, 1, 3):  # vs, vp, rho
            for sgn in (-1, 1):
                m = [a.copy() for a in base]
                v = base[q][i]
                m[q][i] = (v - half * v) if sgn < 0 else (v + half * v)
                models.append(m)
    with open(path, "w") as fh:
        fh.write(f"{len(models)} {L} {periods.size} 1 2 0\n")
        fh.write(" ".join(f"{p:.17g}" for p in periods) + "\n")
        for m in models:
            for a in m:
                fh.write(" ".join(f"{x:.9g}" for x in a) + "\n")
    return len(models)


def run_forward_batch(inputs, nproc):
    """Runs the reference's layered forward driver on every input file, nproc
    processes at a time.  Returns (wall seconds, output paths)."""
    import subprocess
    import time
    if not forward_batch_available():
        raise RuntimeError(f"reference forward driver missing: {FORWARD_BATCH} (make -C oracle ref)")
    outs = [p + ".out" for p in inputs]
    pending = list(zip(inputs, outs))
    running = []
    t0 = time.perf_counter()
    while pending or running:
        while pending and len(running) < nproc:
            fi, fo = pending.pop(0)
            running.append(subprocess.Popen([FORWARD_BATCH, fi, fo], stdout=subprocess.DEVNULL,
                                            stderr=subprocess.DEVNULL))
        p = running.pop(0)
        if p.wait() != 0:
            raise RuntimeError(f"forward driver failed with {p.returncode}")
    return time.perf_counter() - t0, outs


def read_forward_batch(path, nmodels, kmax):
    """Curves (nmodels, kmax) of the driver's output file."""
    c = np.zeros((nmodels, kmax))
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            m, k, _, v = line.split()
            c[int(m) - 1, int(k) - 1] = float(v)
    return c


def ti_column(vs, periods, depz, minthk):
    """Reference TI column (depthkernel_ti_column): pv (kmax,), lsen (kmax, nz-1) float32."""
    if not ref_available():
        raise RuntimeError("reference library not built")
    vs = _f32(vs)
    nz = vs.size
    periods = _f64(periods)
    kmax = periods.size
    pv = np.zeros(kmax, np.float64)
    lsen = np.zeros(kmax * (nz - 1), np.float32)
    i = lambda v: C.byref(C.c_int(v))  # noqa: E731
    ref().depthkernel_ti_column_(i(nz), _ptr(vs), i(kmax), _ptr(periods), i(2), i(0),
                                 _ptr(_f32(depz)), C.byref(C.c_float(minthk)), _ptr(pv),
                                 _ptr(lsen))
    return pv, lsen.reshape(nz - 1, kmax).T.copy()


def ti_at(vs, periods, cp, depz, minthk, layers=False):
    """Reference TI kernels at GIVEN float32 phase velocities cp (kmax,): the model
    of depthkernel_ti_column restated (Brocher and refineLayerMdl through this
    port, moduli TA = rho vp^2, TL = rho vs^2, TF = TA - 2 TL in float32,
    depthkernel_ti_worker.f90:118-142), the reference's own tregn96
    (tregn96.f:52) and the float32 splice restated
    (depthkernel_ti_worker.f90:155-164).  lsen (kmax, nz-1) float32; NaN rows
    where cp = 0 (tregn96 at c = 0), as the reference.  With layers=True also
    returns the chained refined-layer kernels dcR_dA, dcR_dL (kmax, L) float32
    of depthkernelTI.f90:155-156 that the splice sums, and the moduli TA, TL
    (L,) float32 it weights them with."""
    if not ref_available():
        raise RuntimeError("reference library not built")
    vs = _f32(vs)
    nz = vs.size
    vp, rho = brocher(vs)
    depz = _f32(depz)
    out = [np.zeros(NL, np.float32) for _ in range(4)]
    nsub = np.zeros(NL, np.int32)
    L = port().orc_refine(C.c_float(minthk), nz, depz, vp, vs, rho, out[0], out[1], out[2],
                          out[3], _ptr(nsub))
    if L < 0:
        raise ValueError("refined model exceeds NL")
    rvp, rvs, rrho, rthk = out
    f = np.float32
    TA = rrho * (rvp * rvp)
    TL = rrho * (rvs * rvs)
    TF = f(1.0) * (TA - f(2.0) * TL)
    qp = np.full(NL, 150.0, np.float32)
    qs = np.full(NL, 50.0, np.float32)
    zero = np.zeros(NL, np.float32)
    one = np.ones(NL, np.float32)
    kmax = len(periods)
    t_in = _f32(np.asarray(periods, np.float64).astype(np.float32), NP)
    c_in = _f32(np.asarray(cp, np.float32), NP)
    dah, dbv, dn = (np.zeros(NP * NL, np.float32) for _ in range(3))
    i = lambda v: C.byref(C.c_int(v))  # noqa: E731
    ref().tregn96_(i(L), _ptr(rthk), _ptr(TA), _ptr(TA), _ptr(TF), _ptr(TL), _ptr(TL),
                   _ptr(rrho), _ptr(qp), _ptr(qs), _ptr(zero), _ptr(zero), _ptr(one), _ptr(one),
                   i(kmax), _ptr(t_in), _ptr(c_in), _ptr(dah), _ptr(dbv), _ptr(dn))
    dah, dbv, dn = (a.reshape(NL, NP).T[:kmax, :L] for a in (dah, dbv, dn))
    lsen = np.zeros((kmax, nz - 1), np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        den2 = (TA[:L] - f(2.0) * TL[:L]) * (TA[:L] - f(2.0) * TL[:L])
        dA = f(0.5) / (rrho[:L] * rvp[:L]) * dah - TF[:L] / den2 * dn
        dL = f(0.5) / (rrho[:L] * rvs[:L]) * dbv + f(2.0) * TF[:L] / den2 * dn
        tA = dA * TA[:L]
        tL = dL * TL[:L]
        k = 0
        for j in range(nz - 1):
            for _ in range(nsub[j]):
                # left to right, as the reference: (lsen + dA*TA) + dL*TL
                lsen[:, j] = (lsen[:, j] + tA[:, k]) + tL[:, k]
                k += 1
    if layers:
        return lsen, dA, dL, TA[:L].copy(), TL[:L].copy()
    return lsen


# ----------------------------------------------------------------- G rows
def brocher_row_coeffs(v):
    """coe_a, coe_rho of the G builder's Vs row (CalSurfGAniso_Joint.f90:754-760),
    float32 arrays, left to right as the reference's real arithmetic."""
    f = np.float32
    v = np.asarray(v, np.float32)
    v2 = v * v
    v3 = v2 * v
    v4 = v3 * v
    a = f(2.0947) - f(0.8206) * f(2.0) * v + f(0.2683) * f(3.0) * v2 - f(0.0251) * f(4.0) * v3
    p = f(0.9409) + f(2.0947) * v - f(0.8206) * v2 + f(0.2683) * v3 - f(0.0251) * v4
    p2 = p * p
    p3 = p2 * p
    p4 = p3 * p
    r = a * (f(1.6612) - f(0.4721) * f(2.0) * p + f(0.0671) * f(3.0) * p2 -
             f(0.0043) * f(4.0) * p3 + f(0.000106) * f(5.0) * p4)
    return a, r


def g_rows(nvx, nvz, nz, ray_k, fdm, fdmc, fdms, vel, sen_vs, sen_vp, sen_rho, lsen,
           row_base=1, ftol=1.0e-4):
    """The G builder's row fill and compression (CalSurfGAniso_Joint.f90:746-797)
    restated, ray by ray: a dense float32 row of 3*nparpi entries, filled for each
    interior cell (jj, kk) with |fdm(jj,kk)| >= ftol, then (rw, iw, col) of the
    entries with |row| > ftol in increasing col.

    fdm, fdmc, fdms: (nray, nvz+2, nvx+2) float32, [r, jj, kk] = fdm(jj, kk) of ray
    r; ray_k: 0-based period index (knumi - 1); vel (nz, ncol) f32, sen_* (nz, K,
    ncol) f64, lsen (nz-1, K, ncol) f32 with ncol = (nvx+2)*(nvz+2), column
    post = jj*(nvx+2) + kk.  Returns rw f32, iw i32, col i32 (1-based)."""
    f = np.float32
    ftol = f(ftol)
    nd = nz - 1
    cells = nvx * nvz
    nparpi = cells * nd
    nx = nvx + 2
    ca, cr = brocher_row_coeffs(vel[:nd])          # (nd, ncol)
    out_rw, out_iw, out_col = [], [], []
    for r in range(len(ray_k)):
        k = int(ray_k[r])
        row = np.zeros(3 * nparpi, np.float32)
        for jj in range(1, nvz + 1):
            for kk in range(1, nvx + 1):
                w = fdm[r, jj, kk]
                if not abs(w) >= ftol:
                    continue
                post = jj * nx + kk
                z = np.arange(nd)
                idx = z * cells + (jj - 1) * nvx + (kk - 1)
                eff = (sen_vp[:nd, k, post] * ca[:, post].astype(np.float64) +
                       sen_rho[:nd, k, post] * cr[:, post].astype(np.float64)) + sen_vs[:nd, k, post]
                row[idx] = (eff * np.float64(w)).astype(np.float32)
                row[nparpi + idx] = lsen[:, k, post] * fdmc[r, jj, kk]
                row[2 * nparpi + idx] = lsen[:, k, post] * fdms[r, jj, kk]
        keep = np.nonzero(np.abs(row) > ftol)[0]
        out_rw.append(row[keep])
        out_iw.append(np.full(keep.size, row_base + r, np.int32))
        out_col.append((keep + 1).astype(np.int32))
    cat = lambda x, t: np.concatenate(x).astype(t) if x else np.zeros(0, t)  # noqa: E731
    return cat(out_rw, np.float32), cat(out_iw, np.int32), cat(out_col, np.int32)


# ----------------------------------------------------------------- azimuthal map
# the reference's pi: a default-real literal stored in a real*8
# (FwdAzimuthalAniMap.f90:35)
AZ_PI = float(np.float32(3.1415926535898))


_LIBM = None


def _libm_atan2f(y, x):
    """Elementwise atan2f of the C library (the function the flang build calls;
    numpy's float32 arctan2 differs from it in the last bit for some inputs)."""
    global _LIBM
    if _LIBM is None:
        import ctypes.util
        _LIBM = C.CDLL(ctypes.util.find_library("m"))
        _LIBM.atan2f.restype = C.c_float
        _LIBM.atan2f.argtypes = [C.c_float, C.c_float]
    y = np.asarray(y, np.float32)
    x = np.asarray(x, np.float32)
    out = np.empty(y.shape, np.float32)
    fy, fx, fo = y.ravel(), x.ravel(), out.reshape(-1)
    for q in range(fy.size):
        fo[q] = _LIBM.atan2f(float(fy[q]), float(fx[q]))
    return out


def azimuth_map(nvx, nvz, lsen, gcf, gsf, tRcV):
    """FwdAzimuthalAniMap (FwdAzimuthalAniMap.f90:44-89) restated in float32:
    per interior cell and period, cosRc = sum_k Lsen*gcf and sinRc = sum_k Lsen*gsf
    (in k order), AzimAmp = 0.5 sqrt(cos^2 + sin^2), AzimAmpRela = AzimAmp /
    real(tRcV), AzimAng = 0.5 * (atan2(sin, cos)/pi*180, +360 when negative).

    lsen (nz-1, K, ncol) f32 with ncol = (nvx+2)*(nvz+2); gcf, gsf (nz-1, nvz, nvx)
    f32 (the reference's gcf(nx-2, ny-2, nz-1)); tRcV (K, nvx*nvz) f64.  Returns
    dict of (K, nvx*nvz) float32 arrays: cos, sin, amp, amp_rela, angle."""
    f = np.float32
    nd, K, ncol = lsen.shape
    jj, ii = np.meshgrid(np.arange(1, nvz + 1), np.arange(1, nvx + 1), indexing="ij")
    post = (jj * (nvx + 2) + ii).ravel()
    c = np.zeros((K, nvx * nvz), f)
    s = np.zeros((K, nvx * nvz), f)
    for kk in range(nd):
        l = lsen[kk][:, post]
        c = c + l * gcf[kk].ravel()
        s = s + l * gsf[kk].ravel()
    amp = f(0.5) * np.sqrt(c * c + s * s)
    rela = amp / np.asarray(tRcV, np.float64).astype(f)
    ang = (_libm_atan2f(s, c).astype(np.float64) / AZ_PI * 180.0).astype(f)
    ang = np.where(ang < f(0.0), ang + f(360.0), ang)
    ang = f(0.5) * ang
    return dict(cos=c, sin=s, amp=amp, amp_rela=rela, angle=ang)


def ref_azimuth_map(nx, ny, nz, goxd, gozd, dvxd, dvzd, tRc, gcf, gsf, lsen_f, tRcV_f):
    """Runs the reference's FwdAzimuthalAniMap (flang build) in the current
    directory and returns the lines it writes to unit 42 as an (n, 9) array:
    lat, lon, period, isoC, AzimAng, AzimAmpRela, AzimAmp, cos, sin.
    gcf, gsf (nx-2, ny-2, nz-1), lsen_f (nx*ny, K, nz-1), tRcV_f ((nx-2)*(ny-2), K)
    in Fortran order."""
    if not ref_available():
        raise RuntimeError("reference library not built")
    i = lambda v: C.byref(C.c_int(v))  # noqa: E731
    r = lambda v: C.byref(C.c_float(v))  # noqa: E731
    F32 = lambda a: np.asfortranarray(a, np.float32)  # noqa: E731
    if os.path.exists("fort.42"):
        os.remove("fort.42")
    ref().fwdazimuthalanimap_(i(nx), i(ny), i(nz), i(0), r(goxd), r(gozd), r(dvxd), r(dvzd),
                              i(len(tRc)), _ptr(_f64(tRc)), _ptr(F32(gcf)), _ptr(F32(gsf)),
                              _ptr(F32(lsen_f)), _ptr(np.asfortranarray(tRcV_f, np.float64)))
    rows = []
    with open("fort.42") as fh:
        for line in fh:
            line = line.rstrip("\n")
            rows.append([float(line[10 * q:10 * q + 10]) for q in range(9)])
    return np.array(rows)


# ---- first-arrival traveltimes (fmm_port.c, fmm_ref_driver.f90) ----

class FmmDesc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nvx", "nvz", "gdx", "gdz", "sgdl", "sgs", "asgr", "fom")] + \
               [(n, C.c_float) for n in ("goxd", "gozd", "dvxd", "dvzd", "snb", "earth")]


def fmm_desc(nvx, nvz, goxd, gozd, dvxd, dvzd, gdx=5, gdz=5, sgdl=8, sgs=8, asgr=1, fom=1,
             snb=0.5, earth=6371.0):
    """Grid constants of CalSurfGAniso_Joint.f90:375-409 (its defaults)."""
    return FmmDesc(nvx, nvz, gdx, gdz, sgdl, sgs, asgr, fom, goxd, gozd, dvxd, dvzd, snb, earth)


def fmm_grid(d):
    """(nnx, nnz) of the propagation grid."""
    return (d.nvx - 1) * d.gdx + 1, (d.nvz - 1) * d.gdz + 1


def fmm_job(d, pv, scx, scz, rcx, rcz, lib="port"):
    """One (period, source) job: pv the period's (nvz+2)*(nvx+2) phase
    velocities (column fastest along x), source (scx, scz) and receivers in
    radians.  lib "port" (fmm_port.c) or "ref" (the reference's routines).
    Returns dict(rc, ttn (nnx*nnz, Fortran order), times, ttnr, nstsr (refined
    pass, nrx*nrz), nrx, nrz)."""
    nnx, nnz = fmm_grid(d)
    pv = _f64(pv)
    rcx, rcz = _f32(rcx), _f32(rcz)
    nrc = rcx.size
    nmax = ((2 * d.sgs) * d.sgdl + 1) ** 2 + 1
    ttn = np.zeros(nnx * nnz, np.float32)
    ttnr = np.zeros(nmax, np.float32)
    nstsr = np.zeros(nmax, np.int32)
    times = np.zeros(max(nrc, 1), np.float32)
    nrx, nrz, stat = C.c_int(0), C.c_int(0), C.c_int(0)
    if lib == "ref":
        ref().ref_fmm_job(d.nvx, d.nvz, C.c_float(d.goxd), C.c_float(d.gozd), C.c_float(d.dvxd),
                          C.c_float(d.dvzd), d.gdx, d.gdz, d.sgdl, d.sgs, d.asgr, d.fom,
                          C.c_float(d.snb), C.c_float(d.earth), _ptr(pv), C.c_float(scx),
                          C.c_float(scz), nrc, _ptr(rcx), _ptr(rcz), _ptr(ttn), C.byref(nrx),
                          C.byref(nrz), _ptr(ttnr), _ptr(nstsr), _ptr(times), C.byref(stat))
        rc = stat.value
    else:
        L = port()
        L.orc_fmm_job.restype = C.c_int
        rc = L.orc_fmm_job(C.byref(d), _ptr(pv), C.c_float(scx), C.c_float(scz), nrc, _ptr(rcx),
                           _ptr(rcz), _ptr(ttn), C.byref(nrx), C.byref(nrz), _ptr(ttnr),
                           _ptr(nstsr), _ptr(times))
    n = nrx.value * nrz.value
    return dict(rc=rc, ttn=ttn, times=times[:nrc], ttnr=ttnr[:n], nstsr=nstsr[:n],
                nrx=nrx.value, nrz=nrz.value)


def fmm_gridder(d, pv):
    """gridder (CalSurfG.f90:1596-1689) of one period: (nnx*nnz,) float32."""
    nnx, nnz = fmm_grid(d)
    out = np.zeros(nnx * nnz, np.float32)
    port().orc_fmm_gridder(C.byref(d), _ptr(_f64(pv)), _ptr(out))
    return out


def fmm_jobs(d, pv, job_k, scx, scz, rec_off, rcx, rcz, want_ttn=False, nthreads=0):
    """Many jobs (OpenMP over jobs): pv (K, ncol) float64 (pvRc in Fortran
    order), job q at period job_k[q] with receivers rec_off[q]..rec_off[q+1]-1.
    Returns times, ttn ((njob, nnx*nnz) or None), status."""
    L = port()
    L.orc_fmm_jobs.restype = C.c_int
    nnx, nnz = fmm_grid(d)
    job_k = np.ascontiguousarray(job_k, np.int32)
    rec_off = np.ascontiguousarray(rec_off, np.int64)
    njob = job_k.size
    times = np.zeros(max(int(rec_off[-1]), 1), np.float32)
    ttn = np.zeros((njob, nnx * nnz), np.float32) if want_ttn else None
    status = np.zeros(max(njob, 1), np.int32)
    nth = nthreads or L.orc_max_threads()
    L.orc_fmm_jobs(C.byref(d), _ptr(_f64(pv)), njob, _ptr(job_k), _ptr(_f32(scx)),
                   _ptr(_f32(scz)), _ptr(rec_off), _ptr(_f32(rcx)), _ptr(_f32(rcz)), _ptr(times),
                   _ptr(ttn) if want_ttn else None, _ptr(status), nth)
    return times[:int(rec_off[-1])], ttn, status[:njob]
