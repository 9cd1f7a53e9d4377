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
