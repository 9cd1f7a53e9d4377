"""GPU test of the allocation-free root search (dazi_dk_layered_noalloc_dev):
called through the C ABI, its phase velocities, unrounded roots and status
must equal those of dazi_dk_layered_dev bit for bit, on two calls in a row,
the second with a smaller batch (the static work array keeps the first call's
contents, so stale data would show)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def _noalloc(dk, b, with_unrounded, tune=None):
    """Runs the allocation-free entry point on batch b; returns c, c_unrounded
    (or None) and status as host arrays."""
    import torch
    from dazi_current_release_amd import _lib, stage
    s = stage._layered_struct(b)
    c = torch.full((b.nsolves,), -7.0, dtype=torch.float64, device="cuda")
    st = torch.full((b.ncol,), -99, dtype=torch.int32, device="cuda")
    cu = torch.full((b.nsolves,), -7.0, dtype=torch.float64, device="cuda") \
        if with_unrounded else None
    s.c, s.status = c.data_ptr(), st.data_ptr()
    s.c_unrounded = cu.data_ptr() if cu is not None else None
    print(f"noalloc call: ncol={b.ncol} nsolves={b.nsolves} layers={s.nlayers_total}",
          flush=True)
    rc = _lib.lib().dazi_dk_layered_noalloc_dev(
        C.byref(s), None if tune is None else C.byref(tune),
        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    return (c.cpu().numpy(), None if cu is None else cu.cpu().numpy(), st.cpu().numpy())


@pytest.mark.parametrize("tune_args", [None, (32, 4, 0, 64, 2e-6), (16, 4, 1, 64, 2e-6)])
def test_noalloc_equals_layered_dev_twice(dk, tune_args):
    from dazi_current_release_amd import configs
    tune = dk.make_tune(*tune_args) if tune_args else None
    big = configs.layered_workload("ragged_16384", ncol=3000)
    small = configs.layered_workload("ragged_16384", ncol=700, seed=9)
    for W, with_unrounded in ((big, True), (small, False)):
        b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"],
                            W["periods"])
        ref = dk.layered_dev(b, tune=tune)
        want = {k: ref[k].cpu().numpy() for k in ("c", "c_unrounded", "status")}
        c, cu, st = _noalloc(dk, b, with_unrounded, tune)
        assert np.array_equal(st, want["status"])
        assert np.array_equal(c, want["c"])
        if cu is not None:
            assert np.array_equal(cu, want["c_unrounded"])
        assert (c > 0).sum() > 0.9 * c.size


def test_noalloc_float64_models(dk):
    """The differentiable path's float64 models take the same route."""
    from dazi_current_release_amd import configs
    W = configs.layered_workload("diff_4096", ncol=256)
    b = dk.LayeredBatch(W["nl"], W["nk"], *(W[k].astype(np.float64)
                                            for k in ("thk", "vp", "vs", "rho")), W["periods"])
    ref = dk.layered_dev(b)
    c, cu, st = _noalloc(dk, b, True)
    assert np.array_equal(st, ref["status"].cpu().numpy())
    assert np.array_equal(c, ref["c"].cpu().numpy())
    assert np.array_equal(cu, ref["c_unrounded"].cpu().numpy())


def test_noalloc_rejects_oversized_batch(dk):
    import torch
    from dazi_current_release_amd import _lib, configs, stage
    W = configs.layered_workload("ragged_16384", ncol=64)
    b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"], W["periods"])
    s = stage._layered_struct(b)
    c = torch.zeros(b.nsolves, dtype=torch.float64, device="cuda")
    st = torch.zeros(b.ncol, dtype=torch.int32, device="cuda")
    s.c, s.status = c.data_ptr(), st.data_ptr()
    s.nlayers_total = (_lib.NOALLOC_SCRATCH_BYTES // 48) + 1
    rc = _lib.lib().dazi_dk_layered_noalloc_dev(C.byref(s), None, None)
    assert rc == _lib.DAZI_DK_ELIMIT
    assert b"allocation-free" in _lib.lib().dazi_dk_last_error()
