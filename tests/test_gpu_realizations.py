"""GPU tests of the G builder's effective dc/dVs and of the chunked, graph-replayed
forward + backward over Voronoi realizations.

Bars: effective kernel against the reference's expression
(CalSurfGAniso_Joint.f90:754-765: float32 Brocher coefficients, float64
products) within 1e-14 relative to the per-(column, period) max-norm; chunked
graph replay bitwise equal to one eager call over all columns; cell gradients
equal to a host bincount within 1e-12 relative.
"""
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


def coeffs_f32(v):
    """The reference's coefficients in real (float32) arithmetic, left to right."""
    v = np.asarray(v, np.float32)
    f = np.float32
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
    return a.astype(np.float64), r.astype(np.float64)


def test_effective_dvs_matches_reference_expression(dk):
    import torch
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_4096", ncol=256)
    nz = W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"][:, 0, :].T), device="cuda")
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    o = dk.stage_dev(vel, depz, per, W["minthk"], iso=True, ti=False)
    rng = np.random.default_rng(3)
    gc = torch.as_tensor(rng.standard_normal(o["pv"].shape), device="cuda")
    e = dk.effective_dev(vel, o["sen_vs"], o["sen_vp"], o["sen_rho"], gc)
    svs, svp, srho = (o[k].cpu().numpy() for k in ("sen_vs", "sen_vp", "sen_rho"))
    a, r = coeffs_f32(W["vel"][:, 0, :].T)          # (nz, ncol)
    want = svp * a[:, None, :] + srho * r[:, None, :] + svs
    got = e["eff"].cpu().numpy()
    scale = np.abs(want[:-1]).max(axis=0, keepdims=True)
    assert (np.abs(got - want[:-1]) / scale).max() <= 1e-14
    g_want = np.einsum("kc,zkc->zc", gc.cpu().numpy(), want)
    g = e["grad_vs"].cpu().numpy()
    assert np.abs(g - g_want).max() <= 1e-12 * np.abs(g_want).max()


def test_chunked_graph_equals_eager_and_cell_gradient(dk):
    import torch
    from dazi_current_release_amd import configs
    from dazi_current_release_amd.realizations import RealizationStep, cell_gradient
    nreal = 2
    W = configs.voronoi_workload(nreal, ncells=40, nani=20)
    N, nz = W["vel"].shape[0], W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"][:, 0, :].T), device="cuda")
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    T = configs.voronoi_workload(1, seed=11, ncells=40, nani=20)
    vt = torch.as_tensor(np.ascontiguousarray(T["vel"][:, 0, :].T), device="cuda")
    c_true = dk.stage_dev(vt, depz, per, W["minthk"], iso=False, ti=False)["pv"]
    c_obs = c_true.repeat(1, nreal)

    # eager reference over all columns
    o = dk.stage_dev(vel, depz, per, W["minthk"], iso=True, ti=True)
    gc = (o["pv"] - c_obs) * ((o["pv"] > 0) & (c_obs > 0))
    e = dk.effective_dev(vel, o["sen_vs"], o["sen_vp"], o["sen_rho"], gc, eff=False)

    step = RealizationStep(nz, depz, per, W["minthk"], chunk=3000)
    got = {k: torch.zeros_like(v) for k, v in o.items()}
    grad = torch.zeros_like(e["grad_vs"])
    lon = torch.as_tensor(W["layer_of_node"], device="cuda")
    cell = torch.as_tensor(W["cell"], dtype=torch.int64, device="cuda")
    acc = torch.zeros((nreal, 9, 40), dtype=torch.float64, device="cuda")

    def consume(c0, n, s):
        for k in got:
            got[k][..., c0:c0 + n] = s.out[k][..., :n]
        grad[:, c0:c0 + n] = s.grad["grad_vs"][:, :n]
        cell_gradient(s.grad["grad_vs"], lon, cell, 40, c0, n, 4096, acc)

    step.run(vel, c_obs, consume)
    torch.cuda.synchronize()
    for k in o:
        assert torch.equal(got[k], o[k]), k
    assert torch.equal(grad, e["grad_vs"])
    # cell gradient: sum over the nodes of every (realization, layer, cell)
    g = e["grad_vs"].cpu().numpy()
    want = np.zeros((nreal, 9, 40))
    for r in range(nreal):
        for z in range(nz):
            l = W["layer_of_node"][z]
            want[r, l] += np.bincount(W["cell"][r, l], weights=g[z, r * 4096:(r + 1) * 4096],
                                      minlength=40)
    a = acc.cpu().numpy()
    assert np.abs(a - want).max() <= 1e-12 * np.abs(want).max()


def test_cell_sum_chunks_across_realizations_deterministic(dk):
    """dazi_dk_cell_sum_dev through cell_gradient: chunks that start and end
    inside realizations, cell ids outside [0, ncells) skipped, sums equal to a
    float64 bincount and bit-identical between runs."""
    import torch
    from dazi_current_release_amd.realizations import cell_gradient
    rng = np.random.default_rng(5)
    nreal, nl, ncr, ncells, nz = 3, 4, 1000, 37, 9
    lon = torch.as_tensor(np.sort(rng.integers(0, nl, nz)), device="cuda")
    cell_np = rng.integers(0, ncells, (nreal, nl, ncr))
    cell_np[0, 1, :5] = ncells + 3          # outside: skipped
    cell_np[2, 3, 7] = -1
    cell = torch.as_tensor(cell_np, dtype=torch.int64, device="cuda")
    g = rng.standard_normal((nz, nreal * ncr))
    want = np.zeros((nreal, nl, ncells))
    lon_np = lon.cpu().numpy()
    for r in range(nreal):
        for z in range(nz):
            l = lon_np[z]
            ok = (cell_np[r, l] >= 0) & (cell_np[r, l] < ncells)
            want[r, l] += np.bincount(cell_np[r, l][ok], weights=g[z, r * ncr:(r + 1) * ncr][ok],
                                      minlength=ncells)
    runs = []
    for _ in range(2):
        acc = torch.zeros((nreal, nl, ncells), dtype=torch.float64, device="cuda")
        for c0 in range(0, nreal * ncr, 700):
            n = min(700, nreal * ncr - c0)
            gv = torch.as_tensor(np.ascontiguousarray(g[:, c0:c0 + n]), device="cuda")
            cell_gradient(gv, lon, cell, ncells, c0, n, ncr, acc)
        runs.append(acc.cpu().numpy())
    assert np.allclose(runs[0], want, rtol=0, atol=1e-10)
    assert np.array_equal(runs[0], runs[1])
