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
    assert len(table) >= 8 and all(ms > 0 for _, _, ms in table)
    assert (tune.lanes_per_scan, tune.cols_per_wg) == min(table, key=lambda r: r[2])[:2]
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
    """A second depz tensor gives its own geometry; an in-place change needs a
    reset of the cached geometry."""
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
    dk.lib().dazi_dk_reset_geometry()
    c = dk.stage_dev(vel, d1, per, 2.0, iso=False, ti=False)["pv"].clone()
    assert torch.equal(b, c)
