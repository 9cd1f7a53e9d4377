"""GPU tests of the differentiable path: the hand-written VJP kernel
(dazi_dk_layered_vjp_dev) and the torch.autograd.Function over it.

Bars: gradients against central finite differences of the GPU forward with a
relative model step of 1e-4 (float64 models, tight root bracket), within 1e-4
relative; against the contraction sum_k grad_c * dc/dm of the separately
computed refined-layer kernels within 1e-8 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TIGHT = dict(lanes=16, cpw=4, mode=0, tile=64, tol=1e-11)


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def f64_batch(dk, W):
    return dk.LayeredBatch(W["nl"], W["nk"], W["thk"].astype(np.float64),
                           W["vp"].astype(np.float64), W["vs"].astype(np.float64),
                           W["rho"].astype(np.float64), W["periods"])


def tight(dk):
    return dk.make_tune(TIGHT["lanes"], TIGHT["cpw"], TIGHT["mode"], TIGHT["tile"], TIGHT["tol"])


def contraction(b, out, grad_c):
    """sum over each column's periods of grad_c * dc/dm from the kernel pass."""
    nl, nk, loff, toff = b.nl_host, b.nk_host, b.loff_host, b.toff_host
    res = []
    for key in ("dc_dvs", "dc_dvp", "dc_drho"):
        K = out[key].cpu().numpy()
        g = np.zeros(int(nl.sum()))
        for j in range(b.ncol):
            s = slice(toff[j], toff[j] + nk[j])
            g[loff[j]:loff[j] + nl[j]] = K[:nl[j], s] @ grad_c[s]
        res.append(g)
    return res


def test_vjp_equals_kernel_contraction(dk):
    import torch
    from dazi_current_release_amd import configs
    W = configs.layered_workload("diff_4096", ncol=512)
    b = f64_batch(dk, W)
    out = dk.layered_dev(b, kernels=True)
    rng = np.random.default_rng(11)
    gc = rng.standard_normal(b.nsolves)
    gvs, gvp, grho, _ = dk.layered_vjp(b, out["c_unrounded"], torch.as_tensor(gc, device="cuda"))
    ref = contraction(b, out, gc)
    for got, want in zip((gvs, gvp, grho), ref):
        got = got.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max()


def test_vjp_ragged_long_and_empty_columns(dk):
    """Columns with more than 64 periods (several 32-lane chunks), columns with no
    periods (zero gradients) and float32 models."""
    import torch
    from dazi_current_release_amd import configs
    W = configs.layered_workload("ragged_16384", ncol=40)
    rng = np.random.default_rng(2)
    nk = rng.integers(65, 150, size=40).astype(np.int32)
    nk[[3, 17]] = 0
    per = np.concatenate([np.sort(rng.uniform(1.0, 60.0, int(n))) for n in nk])
    b = dk.LayeredBatch(W["nl"], nk, W["thk"], W["vp"], W["vs"], W["rho"], per)
    out = dk.layered_dev(b, kernels=True)
    gc = rng.standard_normal(b.nsolves)
    g = dk.layered_vjp(b, out["c_unrounded"], torch.as_tensor(gc, device="cuda"))[:3]
    ref = contraction(b, out, gc)
    for got, want in zip(g, ref):
        got = got.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max()
        for j in (3, 17):
            assert np.all(got[b.loff_host[j]:b.loff_host[j] + b.nl_host[j]] == 0)


def loss_of(dk, b, vs, vp, rho, w, tune, thk=None):
    out = dk.layered_dev(b, vs=vs, vp=vp, rho=rho, thk=thk, tune=tune)
    return float((out["c_unrounded"].cpu().numpy() * w).sum()), out


@pytest.mark.parametrize("which", ["vs", "vp", "rho", "thk"])
def test_vjp_vs_central_fd_64_columns(dk, which):
    """Directional central differences with a relative model step of 1e-4 on
    64 columns, and per-entry differences on single layers."""
    import torch
    from dazi_current_release_amd import configs
    W = configs.layered_workload("diff_4096", ncol=64)
    b = f64_batch(dk, W)
    tune = tight(dk)
    rng = np.random.default_rng({"vs": 1, "vp": 2, "rho": 3, "thk": 4}[which])
    w = rng.standard_normal(b.nsolves)
    base = {"vs": b.vs, "vp": b.vp, "rho": b.rho, "thk": b.thk}
    _, out0 = loss_of(dk, b, b.vs, b.vp, b.rho, w, tune)
    assert int((out0["c_unrounded"] == 0).sum()) == 0
    g = dict(zip(("vs", "vp", "rho", "thk"),
                 dk.layered_vjp(b, out0["c_unrounded"], torch.as_tensor(w, device="cuda"))))
    grad = g[which].cpu().numpy()
    m0 = base[which]
    h = 1e-4

    def L_at(delta):
        m = m0 + torch.as_tensor(delta, device="cuda")
        args = dict(base)
        args[which] = m.contiguous()
        val, o = loss_of(dk, b, args["vs"], args["vp"], args["rho"], w, tune, thk=args["thk"])
        assert torch.equal(o["status"], out0["status"])
        return val

    mh = m0.cpu().numpy()
    r = rng.uniform(-1.0, 1.0, mh.size)
    if which == "thk":
        # the half-space thickness does not enter: zero gradient, no perturbation
        last = b.loff_host + b.nl_host - 1
        assert np.all(grad[last] == 0.0)
        r[last] = 0.0
    fd = (L_at(h * mh * r) - L_at(-h * mh * r)) / (2 * h)
    an = float((grad * mh * r).sum())
    assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)

    # single entries: top sediment, mid crust, mantle lid, half-space (deepest
    # finite layer for thickness)
    gmax = np.abs(grad).max()
    for j, m in ((0, 0), (5, 12), (33, 30), (60, 39 if which != "thk" else 38)):
        i = int(b.loff_host[j]) + m
        d = np.zeros(mh.size)
        d[i] = h * mh[i]
        fd = (L_at(d) - L_at(-d)) / (2 * d[i])
        assert abs(fd - grad[i]) <= 1e-4 * max(abs(grad[i]), 1e-2 * gmax), (j, m, fd, grad[i])


def test_autograd_checkerboard_misfit(dk):
    """0.5 * sum (c - c_obs)^2 with c_obs from a checkerboard-perturbed model:
    autograd gradient against a directional central difference."""
    import torch
    from dazi_current_release_amd import configs
    from dazi_current_release_amd.autograd import dispersion
    W = configs.layered_workload("diff_4096", ncol=256)
    b = f64_batch(dk, W)
    tune = tight(dk)
    vs_true = configs.checkerboard_vs(W["vs"], W["nl"], ncol_side=16, cell=4).astype(np.float64)
    c_obs = dk.layered_dev(b, vs=torch.as_tensor(vs_true, device="cuda"),
                           tune=tune)["c_unrounded"]
    vs = b.vs.clone().requires_grad_(True)
    vp = b.vp.clone().requires_grad_(True)
    rho = b.rho.clone().requires_grad_(True)
    thk = b.thk.clone().requires_grad_(True)
    c, status = dispersion(vs, vp, rho, b, thk=thk, tune=tune)
    assert int((c == 0).sum()) == 0
    loss = 0.5 * ((c - c_obs) ** 2).sum()
    loss.backward()
    assert vs.grad.dtype == torch.float64 and torch.isfinite(vs.grad).all()

    rng = np.random.default_rng(7)
    h = 1e-4
    dirs = [torch.as_tensor(rng.uniform(-1, 1, b.vs.numel()), device="cuda") for _ in range(4)]

    def L_at(s):
        with torch.no_grad():
            cc, _ = dispersion(b.vs * (1 + s * dirs[0]), b.vp * (1 + s * dirs[1]),
                               b.rho * (1 + s * dirs[2]), b, thk=b.thk * (1 + s * dirs[3]),
                               tune=tune)
            return float(0.5 * ((cc - c_obs) ** 2).sum())

    fd = (L_at(h) - L_at(-h)) / (2 * h)
    an = float((vs.grad * b.vs * dirs[0]).sum() + (vp.grad * b.vp * dirs[1]).sum()
               + (rho.grad * b.rho * dirs[2]).sum() + (thk.grad * b.thk * dirs[3]).sum())
    assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)
