"""Batched execution of the stage: a launch-shape autotuner and HIP-graph replay.

The root search is one launch over the batch, either persistent (mode 0: a
grid that fills the device once, whose groups take the next column of the
batch's scheduling order as they finish) or one group per column (mode 1 with
a tile of the whole batch: the hardware dispatches the workgroups in that
order), with a shape of (lanes per scan step, columns per workgroup).
`autotune` times every admissible (shape, mode) on the caller's own batch with
the library's HIP events and returns the fastest; results do not depend on
the launch beyond float rounding of the root.

`GraphRunner` captures one call of a device entry point (after an identical
warm-up call that sizes the device pool and caches the node geometry, so that
the captured call enqueues kernels only) and replays it: the whole stage
becomes one graph launch.  The reference's counterpart is the pDSurf backend's
per-tile process launch (code/src_inv_iso_joint/CalSurfG.f90:142-280), which
this replaces.
"""
from __future__ import annotations

from . import _lib

# (lanes_per_scan, cols_per_wg): workgroups of 64..256 threads
CANDIDATES = [(8, 8), (8, 16), (8, 32), (16, 4), (16, 8), (16, 16), (32, 2), (32, 4), (32, 8),
              (64, 1), (64, 2), (64, 4)]
# launch modes: 0 persistent, 1 one launch over the whole batch (a tile of
# ONE_LAUNCH columns)
MODES = (0, 1)
ONE_LAUNCH = 1 << 30


def autotune(run, candidates=CANDIDATES, reps=3, refine_tol=2e-6, modes=MODES):
    """run(tune) enqueues one call of the stage with that launch configuration.

    Returns (best tune, table) where table lists (lanes, cols_per_wg, mode,
    root-search ms) for every admissible candidate (shapes whose LDS staging
    does not fit are skipped)."""
    import torch
    lib = _lib.lib()
    table = []
    lib.dazi_dk_enable_timing(1)
    try:
        for mode in modes:
            for lanes, cpw in candidates:
                tune = _lib.make_tune(lanes, cpw, mode, ONE_LAUNCH, refine_tol)
                try:
                    run(tune)
                except _lib.DaziDkError as e:
                    if "LDS" in str(e):
                        continue
                    raise
                best = None
                for _ in range(reps):
                    run(tune)
                    t = _lib.last_timings()[1]
                    best = t if best is None else min(best, t)
                table.append((lanes, cpw, mode, best))
        torch.cuda.synchronize()
    finally:
        lib.dazi_dk_enable_timing(0)
    lanes, cpw, mode, _ = min(table, key=lambda r: r[3])
    return _lib.make_tune(lanes, cpw, mode, ONE_LAUNCH, refine_tol), table


class GraphRunner:
    """fn() enqueues one call of a device entry point on torch's current stream
    with preallocated outputs.  The constructor makes one eager warm-up call,
    then captures a second call in a torch CUDA graph (HIP graph on ROCm);
    replay() launches the graph on the current stream."""

    def __init__(self, fn):
        import torch
        _lib.lib().dazi_dk_enable_timing(0)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()
