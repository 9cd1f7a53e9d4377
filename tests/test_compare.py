"""The A/B dump comparator (host logic): the reference's dump layouts round-trip
and the metrics are what their definitions give on constructed differences."""
import numpy as np

from dazi_current_release_amd import compare


def write_dump(d, arrays):
    d.mkdir(parents=True, exist_ok=True)
    for name, a in arrays.items():
        np.asarray(a).ravel(order="F").tofile(d / name)


def test_metrics_on_constructed_differences(tmp_path):
    rng = np.random.default_rng(0)
    nx, ny, nz, K = 3, 2, 5, 4
    ncol = nx * ny
    base = dict(pv=rng.uniform(3, 4, (ncol, K)),
                **{k: rng.standard_normal((ncol, K, nz)) * 1e-3 for k in ("sen_vs", "sen_vp",
                                                                         "sen_rho")})
    test = {k: v.copy() for k, v in base.items()}
    test["pv"][2, 1] += 5e-6
    test["sen_vs"] *= 1.0 + 1e-3
    lsen = rng.standard_normal((ncol, K, nz - 1)).astype(np.float32)
    fn = {"pv": "pv_depthkernel.bin", "sen_vs": "sen_vs_depthkernel.bin",
          "sen_vp": "sen_vp_depthkernel.bin", "sen_rho": "sen_rho_depthkernel.bin"}
    write_dump(tmp_path / "a", {fn[k]: np.asfortranarray(v) for k, v in base.items()}
               | {"pv_ti.bin": np.asfortranarray(base["pv"]), "lsen_gsc_ti.bin": lsen})
    write_dump(tmp_path / "b", {fn[k]: np.asfortranarray(v) for k, v in test.items()}
               | {"pv_ti.bin": np.asfortranarray(base["pv"]), "lsen_gsc_ti.bin": lsen})
    vel = rng.uniform(3, 4.5, (nx, ny, nz)).astype(np.float32)
    rows = {r["array"]: r for r in compare.compare_dirs(tmp_path / "a", tmp_path / "b", nx, ny,
                                                        nz, K, vel)}
    assert set(rows) == {"pv", "sen_vs", "sen_vp", "sen_rho", "pv_ti", "lsen_gsc",
                         "effective_dcdvs"}
    assert abs(rows["pv"]["max_abs"] - 5e-6) < 1e-12
    assert rows["pv"]["rms_abs"] == np.sqrt((rows["pv"]["max_abs"] ** 2) / (ncol * K))
    assert abs(rows["sen_vs"]["rel_l2"] - 1e-3) < 1e-12
    for k in ("sen_vp", "sen_rho", "pv_ti", "lsen_gsc"):
        assert rows[k]["max_abs"] == 0.0 and rows[k]["rel_l2"] == 0.0
    # crossing counts: a value moved across a threshold is counted once
    assert rows["pv"]["cross_abs_0.0001"] == 0
    # effective kernel: only sen_vs changed, so the difference is 1e-3 * sen_vs
    eff_d = 1e-3 * np.abs(base["sen_vs"]).max()
    assert abs(rows["effective_dcdvs"]["max_abs"] - eff_d) <= 1e-12


def test_cli_writes_csv(tmp_path):
    a = np.arange(12, dtype=np.float64).reshape(6, 2, order="F")
    write_dump(tmp_path / "a", {"pv_depthkernel.bin": a})
    write_dump(tmp_path / "b", {"pv_depthkernel.bin": a + 1.0})
    rows = compare.main(["--base", str(tmp_path / "a"), "--test", str(tmp_path / "b"),
                         "--nx", "3", "--ny", "2", "--nz", "4", "--kmax", "2",
                         "--out", str(tmp_path / "m.csv")])
    assert rows[0]["max_abs"] == 1.0 and rows[0]["mean_abs"] == 1.0
    assert (tmp_path / "m.csv").read_text().startswith("array,")


def test_brocher_coefficients_are_derivatives():
    # coe_a = dVp/dVs and coe_rho = dRho/dVs of the Brocher polynomials
    vs = np.linspace(2.0, 4.5, 11)
    a, r = compare.brocher_coefficients(vs)
    vp = lambda v: 0.9409 + 2.0947 * v - 0.8206 * v**2 + 0.2683 * v**3 - 0.0251 * v**4  # noqa
    rho = lambda p: 1.6612 * p - 0.4721 * p**2 + 0.0671 * p**3 - 0.0043 * p**4 + 0.000106 * p**5  # noqa
    h = 1e-6
    assert np.allclose(a, (vp(vs + h) - vp(vs - h)) / (2 * h), rtol=1e-5)
    assert np.allclose(r, (rho(vp(vs + h)) - rho(vp(vs - h))) / (2 * h), rtol=1e-5)


def test_g_triplet_comparison_sets_and_families(tmp_path):
    """compare_triplets keys by (iw, col): equal runs give zero differences; a
    changed value, a dropped and an extra entry land in the right sets and
    Vs / Gc / Gs families (compare_g_triplet_dumps.py:53-123)."""
    from dazi_current_release_amd.compare import compare_triplets, main
    from dazi_current_release_amd.grows import dump_g_triplets
    nparpi = 10
    iw = np.array([1, 1, 1, 2, 2], np.int32)
    col = np.array([3, 12, 25, 5, 11], np.int32)
    rw = np.array([0.5, -1.0, 2.0, 0.25, 3.0], np.float32)
    rows = {r["set"]: r for r in compare_triplets((iw, col, rw), (iw, col, rw), nparpi)}
    assert rows["common_diff"]["max_abs"] == 0.0 and rows["base_only"]["n"] == 0
    assert (rows["base_all"]["Vs"], rows["base_all"]["Gc"], rows["base_all"]["Gs"]) == (2, 2, 1)
    t_iw = np.array([1, 1, 2, 2, 2], np.int32)
    t_col = np.array([3, 12, 5, 11, 29], np.int32)
    t_rw = np.array([0.5, -1.5, 0.25, 3.0, 1.2e-4], np.float32)
    dump_g_triplets(rw, iw, col, str(tmp_path / "a"))
    dump_g_triplets(t_rw, t_iw, t_col, str(tmp_path / "b"))
    rows = {r["set"]: r for r in main(["--base", str(tmp_path / "a"), "--test",
                                       str(tmp_path / "b"), "--nparpi", str(nparpi)])}
    assert rows["common_diff"]["n"] == 4 and rows["common_diff"]["max_abs"] == 0.5
    assert rows["common_diff_Gc"]["max_abs"] == 0.5 and rows["common_diff_Vs"]["max_abs"] == 0
    assert rows["base_only"]["n"] == 1 and rows["base_only"]["Gs"] == 1
    assert rows["test_only"]["n"] == 1 and rows["test_only"]["abs_le_0.00015"] == 1
