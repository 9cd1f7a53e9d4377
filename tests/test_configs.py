"""CPU tests of the synthetic workload generators."""
import numpy as np

from dazi_current_release_amd import configs
from oracle import port as P


def test_node_workloads_shapes_and_seeds():
    a = configs.node_workload("plain_1024")
    b = configs.node_workload("plain_1024")
    assert a["vel"].shape == (32, 32, 13) and a["vel"].dtype == np.float32
    assert np.array_equal(a["vel"], b["vel"])
    assert len(a["periods"]) == 18 and np.all(np.diff(a["periods"]) > 0)
    assert np.allclose(a["vp"], np.float32(1.73) * a["vel"])
    c = configs.node_workload("plain_4096")
    assert c["vel"].shape == (64, 64, 25) and len(c["periods"]) == 30
    ratio = c["vp"] / c["vel"]
    assert ratio.min() >= 1.70 - 1e-5 and ratio.max() <= 1.80 + 1e-5
    s = configs.node_workload("plain_4096", ncol=10)
    assert s["vel"].shape == (10, 1, 25)
    assert np.array_equal(s["vel"][:, 0], c["vel"].reshape(-1, 25, order="F")[:10])


def test_brocher_matches_oracle_bitwise():
    vs = np.random.default_rng(0).uniform(0.8, 4.8, 1000).astype(np.float32)
    vp, rho = P.brocher(vs)
    assert np.array_equal(configs.brocher_vp(vs), vp)
    assert np.array_equal(configs.brocher_rho(vp), rho)


def test_ragged_workload():
    w = configs.layered_workload("ragged_16384", ncol=200)
    assert w["nl"].min() >= 8 and w["nl"].max() <= 64
    assert w["nk"].min() >= 10 and w["nk"].max() <= 40
    assert w["thk"].size == w["nl"].sum() and w["periods"].size == w["nk"].sum()
    toff = np.concatenate([[0], np.cumsum(w["nk"])])
    for j in range(200):
        p = w["periods"][toff[j]:toff[j + 1]]
        assert np.all(np.diff(p) >= 0) and p.min() >= 1.0 and p.max() <= 60.0
    assert w["vs"].min() >= 0.5


def test_voronoi_workload_projection():
    W = configs.voronoi_workload(2, ncells=10, nani=5)
    N, nz = W["vel"].shape[0], W["vel"].shape[-1]
    assert N == 2 * 4096 and nz == 25
    lon = W["layer_of_node"]
    assert lon[0] == 0 and lon[-1] == 8 and np.all(np.diff(lon) >= 0)
    assert W["cell"].min() >= 0 and W["cell"].max() < 10 and W["ani_cell"].max() < 5
    for r in range(2):
        for z in (0, 7, 24):
            l = lon[z]
            got = W["vel"][r * 4096:(r + 1) * 4096, 0, z]
            assert np.array_equal(got, W["cell_vs"][r, l, W["cell"][r, l]])
    # nearest-nucleus assignment: every column is in some cell and cells are contiguous
    # enough that most cells own a column
    assert len(np.unique(W["cell"][0, 0])) >= 8
    W2 = configs.voronoi_workload(2, ncells=10, nani=5)
    assert np.array_equal(W["vel"], W2["vel"])


def test_cell_gradient_is_a_per_cell_sum():
    import torch
    from dazi_current_release_amd.realizations import cell_gradient
    W = configs.voronoi_workload(2, ncells=10, nani=5)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((25, 2 * 4096))
    acc = torch.zeros((2, 9, 10), dtype=torch.float64)
    lon = torch.as_tensor(W["layer_of_node"])
    cell = torch.as_tensor(W["cell"], dtype=torch.int64)
    for c0 in range(0, 8192, 3000):   # chunks, the last one partial
        n = min(3000, 8192 - c0)
        cell_gradient(torch.as_tensor(g[:, c0:c0 + n]), lon, cell, 10, c0, n, 4096, acc)
    want = np.zeros((2, 9, 10))
    for r in range(2):
        for z in range(25):
            l = W["layer_of_node"][z]
            want[r, l] += np.bincount(W["cell"][r, l], weights=g[z, r * 4096:(r + 1) * 4096],
                                      minlength=10)
    assert np.allclose(acc.numpy(), want, rtol=0, atol=1e-10)
