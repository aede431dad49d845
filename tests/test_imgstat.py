"""The image-parity statistics (tests/imgstat.py) on the CPU: the oracle —
the reference's algorithm restated bit for bit — rendered with another seed
passes them against the reference's high-sample fixtures, and planted
mistakes (a 1 % scale, a mirrored image) fail them."""
import numpy as np
import pytest

import imgstat
import orc
import scenes


def oracle_mean(case, n, seed):
    meta, _, _ = imgstat.load(case)
    o = orc.OracleScene(scenes.load(meta["scene"]), width=meta["width"], height=meta["height"],
                        integrator=meta["integrator"], seed=seed)
    s, c, _ = o.render(0, n)
    return s / c[..., None]


def test_fixture_meta_and_shapes():
    for case in ["c1_path", "c2_path", "c3_path", "c1_direct", "c3_direct", "c2_normals"]:
        meta, m, v = imgstat.load(case)
        assert m.shape == v.shape == (meta["height"], meta["width"], 3)
        assert meta["spp"] >= 4096 and (v >= 0).all()
        assert np.allclose(m.reshape(-1, 3).mean(0), meta["channel_mean"], rtol=1e-6)
        assert meta["stats"]["camera_samples"] == meta["spp"] * meta["width"] * meta["height"]


@pytest.mark.parametrize("case,n,seed", [("c2_path", 256, 11), ("c2_normals", 256, 11)])
def test_oracle_with_another_seed_passes(case, n, seed):
    meta, rm, rv = imgstat.load(case)
    st = imgstat.compare(oracle_mean(case, n, seed), n, rm, rv, meta["spp"])
    imgstat.assert_parity(st)


def test_planted_mistakes_fail():
    meta, rm, rv = imgstat.load("c2_path")
    m = oracle_mean("c2_path", 256, 11)
    scaled = imgstat.compare(m * 1.01, 256, rm, rv, meta["spp"])
    assert min(scaled["channel_z"]) > 3
    mirrored = imgstat.compare(m[:, ::-1], 256, rm, rv, meta["spp"])
    assert mirrored["rmse_ratio"] > 1.3
    shifted = imgstat.compare(np.roll(m, 1, axis=0), 256, rm, rv, meta["spp"])
    assert shifted["rmse_ratio"] > 1.3
    meta_n, rm_n, rv_n = imgstat.load("c2_normals")
    off = imgstat.compare(oracle_mean("c2_normals", 256, 11) + 1e-4, 256, rm_n, rv_n, meta_n["spp"])
    assert not off["flat_ok"]
