"""MI355X-native depth-kernel stage of DAzi surface-wave tomography.

Rayleigh fundamental-mode phase velocities and depth sensitivity kernels per
grid column and period, computed by hand-written HIP kernels for gfx950
behind a C ABI (include/dazi_dk.h).  The Python layer mirrors the reference's
``depthkernel`` interface and exposes device-tensor entry points.
"""
from ._lib import DaziDkError, lib, make_tune, last_timings  # noqa: F401
from .stage import (depthkernel, run_depthkernel, iso_dev, layered_dev, LayeredBatch,  # noqa: F401
                    depthkernelTI, run_depthkernelTI, stage_dev, layered_vjp, effective_dev,
                    node_layers, ti_kernels_dev, ti_vjp_dev)

__all__ = ["DaziDkError", "lib", "make_tune", "last_timings", "depthkernel", "run_depthkernel",
           "iso_dev", "layered_dev", "LayeredBatch", "depthkernelTI", "run_depthkernelTI",
           "stage_dev", "layered_vjp", "effective_dev", "node_layers", "ti_kernels_dev",
           "ti_vjp_dev"]
