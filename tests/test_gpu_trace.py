"""GPU parity of the batch ray queries (rtg_trace) against the oracle.

Model: tests/test_bvh.cpp:34-63 (BVH == brute force on random scenes, with
axis-parallel rays, and any-hit == closest-hit below t_max). Here the GPU
FP64 mode must reproduce the oracle's closest hit exactly — primitive id,
t bit for bit, face flag — and the reference's node-visit / primitive-test
counts; the FP32 mode must agree on all but a small fraction of grazing rays.
"""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu


def ray_set(seed, n=5000, extent=15.0):
    """Origins U(-15,15)^3, unit directions; every 50th ray axis-parallel
    (+z / -z) so inv_dir holds infinities (test_bvh.cpp:42-48)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for i in range(0, n, 50):
        d[i] = (0.0, 0.0, 1.0 if i % 100 else -1.0)
    rays = np.zeros((n, 8))
    rays[:, 0:3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, 1e-6, np.inf
    return rays


@pytest.fixture(scope="module", params=[1, 2, 3])
def random_scene(request, ctx):
    seed = request.param
    print(f"random scene seed {seed}", flush=True)
    o = orc.OracleScene(scenes.random_scene_json(300, 600, seed, seed == 2))
    dev = ctx.scene(o.desc)
    yield seed, o, dev
    dev.release()


def test_f64_closest_matches_oracle_exactly(ctx, random_scene):
    seed, o, dev = random_scene
    rays = ray_set(seed * 100)
    hg, st = ctx.trace(dev, rays, mode=0, use_bvh=1, precision=rtg.PRECISION_F64)
    ho, nv, pt = o.trace(rays, mode=0, use_bvh=1)
    assert np.array_equal(hg["prim"], ho["prim"])
    assert np.array_equal(hg["flags"], ho["flags"])
    assert np.array_equal(hg["t"].view(np.uint64), ho["t"].view(np.uint64))
    assert st.node_visits == nv and st.prim_tests == pt
    assert (ho["flags"] & 1).sum() > 1000  # the rays do hit things


def test_f64_brute_matches_oracle_and_bvh(ctx, random_scene):
    seed, o, dev = random_scene
    rays = ray_set(seed * 100 + 1, n=2000)
    hb, _ = ctx.trace(dev, rays, mode=0, use_bvh=0, precision=rtg.PRECISION_F64)
    ho, _, _ = o.trace(rays, mode=0, use_bvh=0)
    assert np.array_equal(hb["prim"], ho["prim"])
    assert np.array_equal(hb["t"].view(np.uint64), ho["t"].view(np.uint64))
    hv, _ = ctx.trace(dev, rays, mode=0, use_bvh=1, precision=rtg.PRECISION_F64)
    hit = (hb["flags"] & 1) != 0
    assert np.array_equal(hit, (hv["flags"] & 1) != 0)
    np.testing.assert_allclose(hv["t"][hit], hb["t"][hit], rtol=1e-12)


@pytest.mark.parametrize("tmax", [0.5, 5.0, 50.0])
def test_any_hit_agrees_with_closest(ctx, random_scene, tmax):
    seed, o, dev = random_scene
    rays = ray_set(seed * 100 + 2, n=3000)
    closest, _ = ctx.trace(dev, rays, mode=0, precision=rtg.PRECISION_F64)
    r2 = rays.copy()
    r2[:, 7] = tmax
    for prec in (rtg.PRECISION_F64, rtg.PRECISION_F32):
        occ, _ = ctx.trace(dev, r2, mode=1, precision=prec)
        expect = ((closest["flags"] & 1) != 0) & (closest["t"] < tmax)
        got = (occ["flags"] & 1) != 0
        if prec == rtg.PRECISION_F64:
            oo, _, _ = o.trace(r2, mode=1)
            assert np.array_equal(got, (oo["flags"] & 1) != 0)
            assert np.array_equal(got, expect)
        else:
            assert (got != expect).mean() < 2e-3


def test_f32_closest_close_to_oracle(ctx, random_scene):
    seed, o, dev = random_scene
    rays = ray_set(seed * 100 + 3)
    hf, _ = ctx.trace(dev, rays, mode=0, precision=rtg.PRECISION_F32)
    ho, _, _ = o.trace(rays, mode=0)
    same = hf["prim"] == ho["prim"]
    assert same.mean() > 0.998, f"FP32 id mismatches: {(~same).sum()}"
    hit = same & ((ho["flags"] & 1) != 0)
    np.testing.assert_allclose(hf["t"][hit], ho["t"][hit], rtol=2e-4, atol=1e-5)


def test_empty_and_single_primitive_scenes(ctx):
    empty = orc.OracleScene({"materials": {"m": {"type": "lambertian"}}, "objects": []})
    dev = ctx.scene(empty.desc)
    rays = ray_set(5, n=64)
    h, st = ctx.trace(dev, rays, precision=rtg.PRECISION_F64)
    assert (h["flags"] == 0).all() and (h["prim"] == 0xFFFFFFFF).all()
    h32, _ = ctx.trace(dev, rays, precision=rtg.PRECISION_F32)
    assert (h32["flags"] == 0).all()
    dev.release()
    one = orc.OracleScene({"materials": {"m": {"type": "lambertian"}},
                           "objects": [{"type": "sphere", "center": [0, 0, 0], "radius": 2, "material": "m"}]})
    dev = ctx.scene(one.desc)
    r = np.array([[0, 0, 10, 0, 0, -1, 1e-6, np.inf]], dtype=np.float64)
    for prec in (rtg.PRECISION_F64, rtg.PRECISION_F32):
        h, _ = ctx.trace(dev, r, precision=prec)
        assert h["prim"][0] == 0 and abs(h["t"][0] - 8.0) < 1e-5
    dev.release()
