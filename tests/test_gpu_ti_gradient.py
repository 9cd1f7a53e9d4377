"""GPU tests of the TI kernels per refined layer (dc/dA, dc/dL: the values
depthkernelTI splices into Lsen_Gsc, depthkernelTI.f90:152-162), their VJP
kernel and the autograd Function over them.

Bars: dc/dA, dc/dL against the reference's tregn96 and float32 chain at this
build's float32 phase velocities (oracle.port.ti_at) within the TI A/B bands
(run_ti_depthkernel_ab.py:297-304: rel_l2 <= 1e-5 per column, max <= 1e-4 of
the per-(column, period) max-norm); Lsen_Gsc equal bit for bit to the
reference's float32 splice of the same per-layer values; the VJP equal to the
float64 contraction within 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import port as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def dev_node(W):
    import torch
    nz = W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"].reshape(-1, nz, order="F").T),
                          device="cuda")
    return (vel, torch.as_tensor(W["depz"], device="cuda"),
            torch.as_tensor(W["periods"], device="cuda"))


@pytest.mark.parametrize("name", ["plain_4096", "voronoi"])
def test_ti_layer_kernels_vs_reference_chain(dk, name):
    """64 columns: the per-layer kernels against the reference's tregn96 chain,
    and Lsen_Gsc as the reference's float32 splice of this build's values."""
    if not P.ref_available():
        pytest.skip("reference library not built (make -C oracle ref)")
    import torch
    from dazi_current_release_amd import configs
    if name == "voronoi":
        V = configs.voronoi_workload(1)
        cols = np.linspace(0, 4095, 64).astype(int)
        W = dict(vel=V["vel"][cols], depz=V["depz"], periods=V["periods"], minthk=V["minthk"])
    else:
        full = configs.node_workload("plain_4096")
        cols = np.linspace(0, 4095, 64).astype(int)
        v = full["vel"].reshape(-1, full["vel"].shape[-1], order="F")[cols]
        W = dict(vel=v[:, None, :], depz=full["depz"], periods=full["periods"],
                 minthk=full["minthk"])
    vel, depz, per = dev_node(W)
    o = dk.ti_kernels_dev(vel, depz, per, W["minthk"])
    ref_stage = dk.stage_dev(vel, depz, per, W["minthk"], iso=False, ti=True)
    assert torch.equal(o["lsen"], ref_stage["lsen"]) and torch.equal(o["pv"], ref_stage["pv"])
    pv = o["pv"].cpu().numpy()
    kA, kL, ls = (o[k].cpu().numpy() for k in ("dc_dA", "dc_dL", "lsen"))
    L = kA.shape[0]
    f = np.float32
    for j in range(len(cols)):
        vs = W["vel"][j, 0]
        cp = pv[:, j].astype(np.float32)
        rl, rA, rL, TA, TL = P.ti_at(vs, W["periods"], cp, W["depz"], W["minthk"], layers=True)
        ok = cp > 0
        assert rA.shape == (len(cp), L)
        for got, want in ((kA[:, :, j].T, rA), (kL[:, :, j].T, rL)):
            d = (got - want)[ok]
            w = want[ok]
            assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(w), (j, name)
            assert (np.abs(d) / np.abs(w).max(axis=1, keepdims=True)).max() <= 1e-4, (j, name)
        assert np.all(kA[:, ~ok, j] == 0) and np.all(kL[:, ~ok, j] == 0)
        # the reference's float32 splice (depthkernelTI.f90:157) of this build's
        # per-layer values reproduces this build's Lsen_Gsc exactly
        nsub = (L - 1) // (W["vel"].shape[-1] - 1)     # the same for every node interval here
        assert nsub * (W["vel"].shape[-1] - 1) + 1 == L
        for k in np.nonzero(ok)[0]:
            acc = np.zeros(W["vel"].shape[-1] - 1, np.float32)
            m = 0
            for iv in range(acc.size):
                for _ in range(nsub):
                    acc[iv] = (acc[iv] + f(kA[m, k, j]) * TA[m]) + f(kL[m, k, j]) * TL[m]
                    m += 1
            assert np.array_equal(acc, ls[:, k, j]), (j, k)


def test_ti_vjp_matches_contraction(dk):
    import torch
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_4096", ncol=512)
    vel, depz, per = dev_node(W)
    o = dk.ti_kernels_dev(vel, depz, per, W["minthk"], lsen=False)
    rng = np.random.default_rng(4)
    gc = torch.as_tensor(rng.standard_normal(o["pv"].shape), device="cuda")
    gA, gL = dk.ti_vjp_dev(o["dc_dA"], o["dc_dL"], gc)
    g = gc.cpu().numpy()
    for got, k in ((gA, "dc_dA"), (gL, "dc_dL")):
        want = np.einsum("mkc,kc->mc", o[k].cpu().numpy().astype(np.float64), g)
        got = got.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), k


def test_ti_kernels_zero_at_failed_periods(dk):
    """Models whose root search fails part way (the reference's failure
    fixture): no kernel at the failed periods, and the VJP ignores them."""
    import torch
    f = load_golden("ref_ti_failures.npz")
    W = dict(vel=f["vel"][:, None, :], depz=f["depz"], periods=f["periods"],
             minthk=float(f["minthk"]))
    vel, depz, per = dev_node(W)
    o = dk.ti_kernels_dev(vel, depz, per, W["minthk"])
    failed = (o["pv"] == 0)
    assert bool(failed.any())
    for k in ("dc_dA", "dc_dL"):
        assert torch.count_nonzero(o[k][:, failed]) == 0
        assert bool(torch.isfinite(o[k]).all())


def test_ti_dispersion_autograd(dk):
    """TIDispersion: c at zero perturbation is the stage's c; c is linear in
    (dA, dL) with the kernels as slopes; the gradient of a misfit equals the
    contraction of the kernels with its c-gradient."""
    import torch
    from dazi_current_release_amd import configs
    from dazi_current_release_amd.autograd import ti_dispersion
    W = configs.node_workload("plain_4096", ncol=256)
    vel, depz, per = dev_node(W)
    L = dk.node_layers(W["depz"], W["minthk"])
    ncol = vel.shape[1]
    dA = torch.zeros((L, ncol), dtype=torch.float64, device="cuda", requires_grad=True)
    dL = torch.zeros((L, ncol), dtype=torch.float64, device="cuda", requires_grad=True)
    c0, st = ti_dispersion(dA, dL, vel, depz, per, W["minthk"])
    base = dk.stage_dev(vel, depz, per, W["minthk"], iso=False, ti=False)
    assert torch.equal(c0, base["pv"])
    rng = np.random.default_rng(8)
    pA = torch.as_tensor(rng.standard_normal((L, ncol)) * 1e-3, device="cuda")
    pL = torch.as_tensor(rng.standard_normal((L, ncol)) * 1e-3, device="cuda")
    with torch.no_grad():
        c1, _ = ti_dispersion(pA, pL, vel, depz, per, W["minthk"])
    k = dk.ti_kernels_dev(vel, depz, per, W["minthk"], lsen=False)
    lin = (k["dc_dA"].double() * pA[:, None, :]).sum(0) + (k["dc_dL"].double() * pL[:, None, :]).sum(0)
    assert torch.allclose(c1 - c0.detach(), lin, rtol=1e-10, atol=1e-14)
    c_obs = c0.detach() * (1.0 + 0.01 * torch.as_tensor(rng.standard_normal(c0.shape),
                                                          device="cuda"))
    loss = 0.5 * ((c0 - c_obs) ** 2).sum()
    loss.backward()
    r = (c0 - c_obs).detach()
    want_L = (k["dc_dL"].double() * r[None, :, :]).sum(1)
    want_A = (k["dc_dA"].double() * r[None, :, :]).sum(1)
    assert torch.allclose(dL.grad, want_L, rtol=1e-12, atol=1e-18)
    assert torch.allclose(dA.grad, want_A, rtol=1e-12, atol=1e-18)
