"""ctypes binding of libdazi_dk.so (the C ABI in include/dazi_dk.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback in the product path).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdazi_dk.so")

DAZI_DK_OK = 0
DAZI_DK_EINVAL = 9301
DAZI_DK_EHIP = 9302
DAZI_DK_ELIMIT = 9303
DAZI_DK_ENODEV = 9304
NOALLOC_SCRATCH_BYTES = 32 << 20
DEPZ_UNCHANGED = 1
TI_FAILED_NAN = 2

EXPORTS = [
    "dazi_dk_abi_version", "dazi_dk_last_error", "dazi_dk_set_tune", "dazi_dk_get_tune",
    "dazi_dk_iso", "dazi_dk_iso_dev", "dazi_dk_layered_dev", "dazi_dk_last_timings",
    "dazi_dk_enable_timing", "dazi_dk_ti", "dazi_dk_ti_dev", "dazi_dk_stage_dev",
    "dazi_dk_layered_vjp_dev", "dazi_dk_reset_geometry", "dazi_dk_effective_dev",
    "dazi_dk_cell_sum_dev",
    "dazi_dk_pool_bytes", "dazi_dk_grows_dev", "dazi_dk_grows", "dazi_dk_azimuth_map_dev",
    "dazi_dk_azimuth_map", "dazi_dk_travel_plan_create", "dazi_dk_travel_run_dev",
    "dazi_dk_travel_plan_destroy", "dazi_dk_travel", "dazi_dk_layered_noalloc_dev",
    "dazi_dk_node_layers", "dazi_dk_ti_kernels_dev", "dazi_dk_ti_vjp_dev",
]


class DaziDkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dazi_dk error {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("kmax", C.c_int32),
                ("iwave", C.c_int32), ("igr", C.c_int32), ("tile_columns", C.c_int32),
                ("vp_rho_mode", C.c_int32), ("minthk0", C.c_float), ("flags", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("lanes_per_scan", C.c_int32), ("cols_per_wg", C.c_int32), ("mode", C.c_int32),
                ("tile_columns", C.c_int32), ("refine_tol", C.c_double)]


class Layered(C.Structure):
    _fields_ = [("ncol", C.c_int32), ("max_layers", C.c_int32), ("nsolves", C.c_int64),
                ("nl", C.c_void_p), ("loff", C.c_void_p), ("nk", C.c_void_p),
                ("toff", C.c_void_p), ("solve_col", C.c_void_p), ("thk", C.c_void_p),
                ("vp", C.c_void_p), ("vs", C.c_void_p), ("rho", C.c_void_p),
                ("periods", C.c_void_p), ("c", C.c_void_p), ("status", C.c_void_p),
                ("dc_dvs", C.c_void_p), ("dc_dvp", C.c_void_p), ("dc_drho", C.c_void_p),
                ("c_unrounded", C.c_void_p), ("thk64", C.c_void_p), ("vp64", C.c_void_p),
                ("vs64", C.c_void_p), ("rho64", C.c_void_p), ("nlayers_total", C.c_int64)]


class AzimDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("kmax", C.c_int32)]


class GrowDesc(C.Structure):
    _fields_ = [("nvx", C.c_int32), ("nvz", C.c_int32), ("nz", C.c_int32), ("kmax", C.c_int32),
                ("nray", C.c_int64), ("row_base", C.c_int32), ("ftol", C.c_float)]


class TravelDesc(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nvx", "nvz", "gdx", "gdz", "sgdl", "sgs", "asgr",
                                         "fom")] + \
               [(n, C.c_float) for n in ("goxd", "gozd", "dvxd", "dvzd", "snb", "earth")] + \
               [("kmax", C.c_int32), ("slots", C.c_int32)]


_lib = None


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c \"import __graft_entry__ as g; "
                f"g.build()\"` (the stage has no CPU fallback)")
        # PyTorch-ROCm ships its own libamdhip64.so.7.  Loading torch first makes
        # this library bind to that same runtime (same soname), so a process
        # that uses both has one HIP runtime instead of two competing ones.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.dazi_dk_abi_version.restype = C.c_int32
        L.dazi_dk_last_error.restype = C.c_char_p
        L.dazi_dk_set_tune.argtypes = [C.POINTER(Tune)]
        L.dazi_dk_get_tune.argtypes = [C.POINTER(Tune)]
        L.dazi_dk_iso.argtypes = [C.POINTER(Desc)] + [C.c_void_p] * 9
        L.dazi_dk_iso_dev.argtypes = [C.POINTER(Desc), C.c_void_p] + [C.c_void_p] * 11
        L.dazi_dk_layered_dev.argtypes = [C.POINTER(Layered), C.c_void_p, C.c_void_p]
        L.dazi_dk_layered_noalloc_dev.argtypes = [C.POINTER(Layered), C.c_void_p, C.c_void_p]
        L.dazi_dk_ti.argtypes = [C.POINTER(Desc)] + [C.c_void_p] * 5
        L.dazi_dk_ti_dev.argtypes = [C.POINTER(Desc), C.c_void_p] + [C.c_void_p] * 9
        L.dazi_dk_stage_dev.argtypes = [C.POINTER(Desc), C.c_void_p] + [C.c_void_p] * 12
        L.dazi_dk_layered_vjp_dev.argtypes = [C.POINTER(Layered)] + [C.c_void_p] * 6
        L.dazi_dk_node_layers.argtypes = [C.c_int32, C.c_void_p, C.c_float,
                                          C.POINTER(C.c_int32)]
        L.dazi_dk_ti_kernels_dev.argtypes = [C.POINTER(Desc), C.c_void_p] + [C.c_void_p] * 11
        L.dazi_dk_ti_vjp_dev.argtypes = [C.c_int32] * 3 + [C.c_void_p] * 6
        L.dazi_dk_last_timings.argtypes = [C.POINTER(C.c_double), C.c_int32]
        L.dazi_dk_last_timings.restype = C.c_int32
        L.dazi_dk_enable_timing.argtypes = [C.c_int32]
        L.dazi_dk_reset_geometry.argtypes = []
        L.dazi_dk_effective_dev.argtypes = [C.POINTER(Desc)] + [C.c_void_p] * 8
        L.dazi_dk_cell_sum_dev.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int64,
                                           C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                           C.c_int32, C.c_void_p, C.c_void_p]
        L.dazi_dk_pool_bytes.argtypes = [C.POINTER(C.c_int64)]
        L.dazi_dk_grows_dev.argtypes = ([C.POINTER(GrowDesc)] + [C.c_void_p] * 11 +
                                        [C.c_int64] + [C.c_void_p] * 6)
        L.dazi_dk_grows.argtypes = ([C.POINTER(GrowDesc)] + [C.c_void_p] * 11 + [C.c_int64] +
                                    [C.c_void_p] * 4)
        L.dazi_dk_azimuth_map_dev.argtypes = [C.POINTER(AzimDesc)] + [C.c_void_p] * 10
        L.dazi_dk_azimuth_map.argtypes = [C.POINTER(AzimDesc)] + [C.c_void_p] * 9
        L.dazi_dk_travel_plan_create.argtypes = ([C.POINTER(TravelDesc), C.c_int64] +
                                                 [C.c_void_p] * 6 + [C.POINTER(C.c_void_p)])
        L.dazi_dk_travel_run_dev.argtypes = [C.c_void_p] * 6
        L.dazi_dk_travel_plan_destroy.argtypes = [C.c_void_p]
        L.dazi_dk_travel.argtypes = [C.POINTER(TravelDesc), C.c_int64] + [C.c_void_p] * 10
        for name in EXPORTS:
            if not hasattr(L, name):
                raise RuntimeError(f"{LIB_PATH} does not export {name}")
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().dazi_dk_last_error()
        raise DaziDkError(rc, msg.decode() if msg else "")


def make_tune(lanes_per_scan=0, cols_per_wg=0, mode=0, tile_columns=0, refine_tol=0.0) -> Tune:
    return Tune(lanes_per_scan, cols_per_wg, mode, tile_columns, refine_tol)


def last_timings():
    """[prep, root search, iso kernels, TI kernels] in ms, the root-search launch
    count, the refined layers per column and the TI launch's device threads."""
    buf = (C.c_double * 7)()
    n = lib().dazi_dk_last_timings(buf, 7)
    return list(buf)[:n]
