"""GPU tests of the G-row assembly (dazi_dk_grows_dev / dazi_dk_grows) against
the oracle's restatement of the reference's row fill and compression
(CalSurfGAniso_Joint.f90:746-797).

Inputs: the GPU stage's own outputs (sen_vs, sen_vp, sen_rho, Lsen_Gsc) for a
32x32-column grid, and synthetic straight rays.  Bar: bit-exact triplets (the
values are float32 products of the same float64/float32 operands, evaluated
without contraction; the filter and the order are exact).
"""
import numpy as np
import pytest

from oracle import port as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    from dazi_current_release_amd import configs
    m.lib()
    W = configs.node_workload("plain_1024")
    nz = W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"].reshape(-1, nz, order="F").T),
                          device="cuda")
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    o = m.stage_dev(vel, depz, per, W["minthk"], iso=True, ti=True)
    torch.cuda.synchronize()
    R = configs.ray_workload(30, 30, len(W["periods"]), 6, seed=4)
    host = {k: o[k].cpu().numpy() for k in ("sen_vs", "sen_vp", "sen_rho", "lsen")}
    return m, W, vel, o, R, host


def oracle_rows(R, host, vel_np, rays=None, row_base=1):
    from dazi_current_release_amd import configs
    rays = range(len(R["ray_k"])) if rays is None else rays
    dense = [configs.ray_dense(R, r) for r in rays]
    fdm, fdmc, fdms = (np.array([m[i] for m in dense]) for i in range(3))
    k = R["ray_k"][list(rays)]
    nz = host["sen_vs"].shape[0]
    return P.g_rows(30, 30, nz, k, fdm, fdmc, fdms, vel_np, host["sen_vs"], host["sen_vp"],
                    host["sen_rho"], host["lsen"], row_base=row_base)


def test_device_rows_equal_the_oracle_bit_for_bit(setup):
    from dazi_current_release_amd import grows
    m, W, vel, o, R, host = setup
    g = grows.g_rows_dev(grows.rays_to_device(R), 30, 30, vel, o["sen_vs"], o["sen_vp"],
                         o["sen_rho"], o["lsen"])
    n = int(g["nar"].item())
    rw, iw, col = (g[k][:n].cpu().numpy() for k in ("rw", "iw", "col"))
    want = oracle_rows(R, host, vel.cpu().numpy())
    print("rays", len(R["ray_k"]), "entries", n)
    assert n == want[0].size > 0
    assert np.array_equal(iw, want[1])
    assert np.array_equal(col, want[2])
    assert np.array_equal(rw, want[0])
    cnt = np.bincount(iw - 1, minlength=len(R["ray_k"]))
    assert np.array_equal(g["row_nnz"].cpu().numpy(), cnt)
    info = grows.check_triplets(iw, col, rw, len(R["ray_k"]), 3 * 30 * 30 * 12)
    assert info["breaks"] == 0


def test_host_entry_equals_device_entry_and_dumps(setup, tmp_path, monkeypatch):
    from dazi_current_release_amd import grows
    m, W, vel, o, R, host = setup
    monkeypatch.setenv("DAZI_G_TRIPLET_DUMP_DIR", str(tmp_path / "g"))
    rw, iw, col = grows.g_rows(R, 30, 30, vel.cpu().numpy(), host["sen_vs"], host["sen_vp"],
                               host["sen_rho"], host["lsen"], row_base=11)
    want = oracle_rows(R, host, vel.cpu().numpy(), row_base=11)
    assert np.array_equal(rw, want[0]) and np.array_equal(iw, want[1])
    assert np.array_equal(col, want[2])
    assert np.array_equal(np.fromfile(tmp_path / "g" / "rw_triplets.bin", np.float32), rw)
    assert np.array_equal(np.fromfile(tmp_path / "g" / "col_triplets.bin", np.int32), col)


def test_cap_empty_and_subset_batches(setup):
    import torch
    from dazi_current_release_amd import grows
    m, W, vel, o, R, host = setup
    rays = grows.rays_to_device(R)
    full = grows.g_rows_dev(rays, 30, 30, vel, o["sen_vs"], o["sen_vp"], o["sen_rho"],
                            o["lsen"])
    n = int(full["nar"].item())
    # a cap below the count: the first cap entries, and nar still the full count
    cap = n // 3
    out = dict(rw=torch.zeros(cap, dtype=torch.float32, device="cuda"),
               iw=torch.zeros(cap, dtype=torch.int32, device="cuda"),
               col=torch.zeros(cap, dtype=torch.int32, device="cuda"))
    part = grows.g_rows_dev(rays, 30, 30, vel, o["sen_vs"], o["sen_vp"], o["sen_rho"], o["lsen"],
                            out=out)
    torch.cuda.synchronize()
    assert int(part["nar"].item()) == n
    for k in ("rw", "iw", "col"):
        assert torch.equal(part[k], full[k][:cap])
    # empty batch
    empty = {k: v[:0] for k, v in rays.items()}
    empty["cell_off"] = rays["cell_off"][:1]
    e = grows.g_rows_dev(empty, 30, 30, vel, o["sen_vs"], o["sen_vp"], o["sen_rho"], o["lsen"])
    assert int(e["nar"].item()) == 0
    # rays 5..9 alone, numbered from row 6: the same rows as in the full batch
    sub = dict(R)
    o0, o1 = R["cell_off"][5], R["cell_off"][10]
    sub = dict(ray_k=R["ray_k"][5:10], cell_off=R["cell_off"][5:11] - o0,
               cell=R["cell"][o0:o1], fdm=R["fdm"][o0:o1], fdmc=R["fdmc"][o0:o1],
               fdms=R["fdms"][o0:o1])
    s = grows.g_rows_dev(grows.rays_to_device(sub), 30, 30, vel, o["sen_vs"], o["sen_vp"],
                         o["sen_rho"], o["lsen"], row_base=6)
    ns = int(s["nar"].item())
    fiw = full["iw"][:n]
    sel = (fiw >= 6) & (fiw <= 10)
    for k in ("rw", "iw", "col"):
        assert torch.equal(s[k][:ns], full[k][:n][sel])


@pytest.mark.parametrize("nz", [7, 9, 70])
def test_scalar_path_long_rays_and_other_depth_counts(setup, nz):
    """Rays of more than 128 cells (a 150-cell-wide grid) and depth counts that
    are not a multiple of 4 (nz = 7, 70) take the scalar-load path; nz = 9 mixes
    vector rays with long ones; nz = 70 also takes the untiled planes kernel."""
    import torch
    from dazi_current_release_amd import configs, grows
    nvx, nvz, K = 150, 6, 3
    rng = np.random.default_rng(nz)
    ncol = (nvx + 2) * (nvz + 2)
    vel = rng.uniform(2.5, 4.5, (nz, ncol)).astype(np.float32)
    sv = [rng.standard_normal((nz, K, ncol)) * 0.05 for _ in range(3)]
    ls = (rng.standard_normal((nz - 1, K, ncol)) * 0.05).astype(np.float32)
    R = configs.ray_workload(nvx, nvz, K, 5, seed=7, cell_km=0.5)  # a 138-cell ray
    ncell = np.diff(R["cell_off"])
    print("cells per ray: max", ncell.max(), "min", ncell.min())
    assert ncell.max() > 128
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    g = grows.g_rows_dev(grows.rays_to_device(R), nvx, nvz, T(vel), T(sv[0]), T(sv[1]),
                         T(sv[2]), T(ls))
    n = int(g["nar"].item())
    rw, iw, col = (g[k][:n].cpu().numpy() for k in ("rw", "iw", "col"))
    dense = [configs.ray_dense(R, r) for r in range(len(R["ray_k"]))]
    fdm, fdmc, fdms = (np.array([m[i] for m in dense]) for i in range(3))
    want = P.g_rows(nvx, nvz, nz, R["ray_k"], fdm, fdmc, fdms, vel, sv[0], sv[1], sv[2], ls)
    assert n == want[0].size > 0
    assert np.array_equal(iw, want[1]) and np.array_equal(col, want[2])
    assert np.array_equal(rw, want[0])
