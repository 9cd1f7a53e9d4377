"""GPU tests of the batched execution: the launch-shape autotuner and HIP-graph
replay of whole stage calls (ragged layered batches and node grids).

Every graph replay is compared with an eager call of the same entry point, with
eager work and synchronisations between replays; bars: bitwise equality (the
same kernels on the same inputs)."""
import numpy as np
import pytest

from oracle import port as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def test_autotune_ragged_picks_a_valid_shape(dk):
    from dazi_current_release_amd import batched, configs
    W = configs.layered_workload("ragged_16384", ncol=2000)
    b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"], W["periods"])
    out = b.alloc_outputs(False)
    tune, table = batched.autotune(lambda t: dk.layered_dev(b, tune=t, out=out))
    assert len(table) >= 16 and all(ms > 0 for *_, ms in table)
    assert {m for _, _, m, _ in table} == {0, 1}
    assert (tune.lanes_per_scan, tune.cols_per_wg, tune.mode) == min(table,
                                                                     key=lambda r: r[3])[:3]
    o = {k: v.cpu().numpy() for k, v in dk.layered_dev(b, tune=tune).items()}
    ref, fail = P.forward_layered(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"],
                                  W["periods"], nthreads=16)
    assert np.array_equal(o["c"] == 0, ref == 0)
    assert np.abs(o["c"] - ref).max() <= 1e-5
    assert np.array_equal(o["status"], fail)


def test_graph_replay_layered_with_kernels(dk):
    import torch
    from dazi_current_release_amd import batched, configs
    W = configs.layered_workload("ragged_16384", ncol=3000)
    b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"], W["periods"])
    tune = dk.make_tune(16, 16, 0, 64, 2e-6)
    eager = {k: v.clone() for k, v in dk.layered_dev(b, kernels=True, tune=tune).items()}
    out = b.alloc_outputs(True)
    g = batched.GraphRunner(lambda: dk.layered_dev(b, kernels=True, tune=tune, out=out))
    other = configs.layered_workload("ragged_16384", seed=9, ncol=500)
    b2 = dk.LayeredBatch(other["nl"], other["nk"], other["thk"], other["vp"], other["vs"],
                         other["rho"], other["periods"])
    for rep in range(3):
        for v in out.values():
            v.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(out[k], eager[k]), (rep, k)
        dk.layered_dev(b2, kernels=True)  # eager work between replays
        torch.cuda.synchronize()


def test_graph_replay_stage_iso_and_ti(dk):
    import torch
    from dazi_current_release_amd import batched, configs
    W = configs.node_workload("plain_4096", ncol=1024)
    nz = W["vel"].shape[-1]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a.reshape(-1, nz, order="F").T),  # noqa
                                  device="cuda")
    vel, vp, rho = T(W["vel"]), T(W["vp"]), T(W["rho"])
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    run = lambda out=None: dk.stage_dev(vel, depz, per, W["minthk"], vp=vp, rho=rho,  # noqa
                                        iso=True, ti=True, out=out)
    eager = {k: v.clone() for k, v in run().items()}
    out = {k: torch.empty_like(v) for k, v in eager.items()}
    g = batched.GraphRunner(lambda: run(out))
    for rep in range(3):
        for v in out.values():
            v.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(out[k], eager[k]), (rep, k)
        # eager work between replays; from the second replay on with another node
        # geometry (other depths and sublayer count), which must not touch the
        # geometry the graph was captured with
        if rep == 0:
            dk.stage_dev(vel[:, :100].contiguous(), depz, per, W["minthk"], iso=True, ti=False)
        else:
            dk.stage_dev(vel[:, :100].contiguous(), depz * (1.0 + 0.25 * rep), per,
                         W["minthk"] + rep, iso=True, ti=True)
        torch.cuda.synchronize()


def test_geometry_cache_follows_depz(dk):
    """A second depz tensor gives its own geometry; an in-place change of depz
    is seen without any reset."""
    import torch
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_1024", ncol=64)
    nz = W["vel"].shape[-1]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a.reshape(-1, nz, order="F").T),  # noqa
                                  device="cuda")
    vel = T(W["vel"])
    per = torch.as_tensor(W["periods"], device="cuda")
    d1 = torch.as_tensor(W["depz"], device="cuda")
    a = dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"].clone()
    d2 = d1 * 1.5
    b = dk.stage_dev(vel, d2, per, 2.0, iso=False, ti=False)["pv"].clone()
    assert not torch.equal(a, b)
    # back to the first geometry, then another sublayer count on it
    assert torch.equal(a, dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"])
    e = dk.stage_dev(vel, d1, per, 3.0, iso=False, ti=False)["pv"].clone()
    assert not torch.equal(a, e)
    assert torch.equal(a, dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"])
    d1.mul_(1.5)
    c = dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"].clone()
    assert torch.equal(b, c)


def _geometry_case(dk):
    import torch
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_1024", ncol=64)
    nz = W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"].reshape(-1, nz, order="F").T),
                          device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    return W, vel, per


def test_geometry_cache_keyed_on_depz_values_not_address(dk):
    """depz freed and another depz with other depths allocated at the same
    address: the cached geometry must not be reused (no reset call), through
    the Python layer and through the raw C ABI without the unchanged flag."""
    import ctypes as C
    import torch
    from dazi_current_release_amd import _lib
    W, vel, per = _geometry_case(dk)
    want = {}
    for scale in (1.0, 1.5):
        d = torch.as_tensor(W["depz"] * np.float32(scale), device="cuda")
        want[scale] = dk.stage_dev(vel, d, per, 2.0, iso=False, ti=False)["pv"].clone()
        del d
    assert not torch.equal(want[1.0], want[1.5])
    # a block of memory that holds one depz, then (after that tensor is gone)
    # another depz with other depths at the same address: what the caching
    # allocator does when a freed block is handed out again
    nz = len(W["depz"])
    block = torch.empty(1024, dtype=torch.float32, device="cuda")
    d1 = block[:nz]
    d1.copy_(torch.as_tensor(W["depz"]))
    addr = d1.data_ptr()
    a = dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"].clone()
    assert torch.equal(a, want[1.0])
    del d1
    block[:nz].copy_(torch.as_tensor(W["depz"] * np.float32(1.5)))
    d2 = block[:nz]
    assert d2.data_ptr() == addr
    b = dk.stage_dev(vel, d2, per, 2.0, iso=False, ti=False)["pv"].clone()
    assert torch.equal(b, want[1.5])
    # the raw ABI with flags = 0 (no promise that depz is unchanged): new
    # values at the same address are seen, since the library compares them
    pv = torch.empty_like(b)
    st = torch.empty(vel.shape[1], dtype=torch.int32, device="cuda")
    desc = _lib.Desc(vel.shape[1], 1, nz, per.numel(), 2, 0, 64, 0, 2.0, 0)
    d2.copy_(torch.as_tensor(W["depz"], device="cuda"))
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _lib.lib().dazi_dk_iso_dev(C.byref(desc), None, C.c_void_p(vel.data_ptr()), None, None,
                                    C.c_void_p(d2.data_ptr()), C.c_void_p(per.data_ptr()),
                                    C.c_void_p(pv.data_ptr()), None, None, None,
                                    C.c_void_p(st.data_ptr()), s)
    _lib.check(rc)
    torch.cuda.synchronize()
    assert torch.equal(pv, want[1.0])


def test_graph_replay_with_changed_depz_reports_geometry_status(dk):
    """A captured call keeps its geometry; a replay after depz changed gives
    status -2 and zero phase velocities instead of stale results."""
    import torch
    from dazi_current_release_amd import batched
    W, vel, per = _geometry_case(dk)
    depz = torch.as_tensor(W["depz"], device="cuda")
    out = dk.stage_dev(vel, depz, per, 2.0, iso=True, ti=True)
    want = {k: v.clone() for k, v in out.items()}
    g = batched.GraphRunner(lambda: dk.stage_dev(vel, depz, per, 2.0, iso=True, ti=True,
                                                 out=out))
    g.replay()
    torch.cuda.synchronize()
    for k in want:
        assert torch.equal(out[k], want[k]), k
    orig = depz.clone()
    depz.mul_(1.25)
    g.replay()
    torch.cuda.synchronize()
    assert torch.all(out["status"] == -2)
    for k in ("pv", "sen_vs", "sen_vp", "sen_rho", "lsen"):
        assert torch.count_nonzero(out[k]) == 0, k
    depz.copy_(orig)
    g.replay()
    torch.cuda.synchronize()
    for k in want:
        assert torch.equal(out[k], want[k]), k


def test_launch_mode_and_shape_do_not_change_results(dk):
    """At one lanes-per-scan value every column runs the same sequence of
    operations whatever the launch: persistent or one launch over the batch,
    any columns per workgroup, tiles of any size.  Roots, status and
    refined-layer kernels are bit-identical."""
    from dazi_current_release_amd import batched, configs
    W = configs.layered_workload("ragged_16384", ncol=3000, seed=12)
    b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"], W["periods"])
    runs = {}
    for lanes, cpw, mode, tile in ((16, 4, 0, 64), (16, 16, 0, 64), (16, 4, 1, batched.ONE_LAUNCH),
                                   (16, 8, 1, 257)):
        o = dk.layered_dev(b, kernels=True, tune=dk.make_tune(lanes, cpw, mode, tile, 2e-6))
        runs[(lanes, cpw, mode, tile)] = {k: v.cpu().numpy() for k, v in o.items()}
    ref = runs[(16, 4, 0, 64)]
    for key, o in runs.items():
        for k in ("c", "c_unrounded", "status", "dc_dvs", "dc_dvp", "dc_drho"):
            assert np.array_equal(o[k], ref[k]), (key, k)
