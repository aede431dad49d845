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
    print(f"random scene {seed}: FP32 id mismatches {(~same).sum()} of {len(rays)}", flush=True)
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


def cornell_surface_rays(o, n, seed):
    """Rays leaving the surfaces of the 555-unit Cornell box as the path
    integrator spawns them: a uniform point on a random triangle, offset
    along the geometric normal by the reference's spawn offset
    (include/raytracer/math/ray.hpp:42-58: 1.6e-6 max(1, |p|_inf)) to the
    side the direction leaves on, direction uniform over that hemisphere.
    Components are rounded to float, so both precisions trace the same
    rays."""
    prims = orc.arrays(o.desc)["prims"].view(np.float64).reshape(-1, 20)
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(prims), n)
    v0, v1, v2 = prims[idx, 2:5], prims[idx, 5:8], prims[idx, 8:11]
    a, b = rng.random(n), rng.random(n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = v0 + (v1 - v0) * a[:, None] + (v2 - v0) * b[:, None]
    gn = np.cross(v1 - v0, v2 - v0)
    gn /= np.linalg.norm(gn, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    side = np.sign((d * gn).sum(1))[:, None]
    side[side == 0] = 1
    orig = p + gn * side * (1.6e-6 * np.maximum(1.0, np.abs(p).max(1)))[:, None]
    rays = np.zeros((n, 8))
    rays[:, 0:3] = orig.astype(np.float32)
    rays[:, 3:6] = d.astype(np.float32)
    rays[:, 6], rays[:, 7] = 1e-6, np.inf
    return rays


def f32_walk_vs_brute(ctx, dev, rays):
    """FP32 BVH walk against FP32 brute force over every primitive (the same
    slot test on the same float data): (rays whose t or hit flag differ,
    ids that differ at equal t)."""
    hb, _ = ctx.trace(dev, rays, mode=0, use_bvh=1, precision=rtg.PRECISION_F32)
    hr, _ = ctx.trace(dev, rays, mode=0, use_bvh=0, precision=rtg.PRECISION_F32)
    differ = (hb["t"].view(np.uint64) != hr["t"].view(np.uint64)) | (hb["flags"] != hr["flags"])
    return int(differ.sum()), int((hb["prim"] != hr["prim"]).sum())


def test_f32_walk_culls_no_box_on_random_scenes(ctx, random_scene):
    """The FP32 walk's closest t equals FP32 brute force's bit for bit on
    every ray: a box culled by slab rounding would show as a farther t or a
    miss."""
    seed, o, dev = random_scene
    rays = ray_set(seed * 100 + 5, n=1 << 20)
    rays[:, :6] = rays[:, :6].astype(np.float32)
    differ, ties = f32_walk_vs_brute(ctx, dev, rays)
    assert differ == 0 and ties == 0


def test_f32_walk_culls_no_box_far_from_origin(ctx):
    """The same 1e4 units from the world origin. The node records hold half
    floats in a frame centred on the scene and a ray takes its slab offsets
    from that centre, so the boxes stay tight there; the FP32 walk's closest
    t must still equal FP32 brute force's bit for bit on every ray (ids may
    differ only between primitives hit at exactly the same t)."""
    off = np.array([1.0e4, -3.0e3, 2.5e3])
    o = orc.OracleScene(scenes.translated(scenes.random_scene_json(300, 600, 1, False), off))
    dev = ctx.scene(o.desc)
    try:
        rays = ray_set(177, n=1 << 20)
        rays[:, 0:3] += off
        rays[:, :6] = rays[:, :6].astype(np.float32)
        differ, ties = f32_walk_vs_brute(ctx, dev, rays)
        h, st = ctx.trace(dev, rays, mode=0, use_bvh=1, precision=rtg.PRECISION_F32)
        print(f"far scene: {int((h['flags'] & 1).sum())} hits, {st.node_visits / len(rays):.2f} node boxes per ray; "
              f"walk vs brute: {differ} differ, {ties} ties", flush=True)
        assert int((h["flags"] & 1).sum()) > 100000
        assert differ == 0 and ties <= 1e-4 * len(rays)
    finally:
        dev.release()


def test_f32_cornell_scale_rays(ctx):
    """At the 555-unit scale of the Cornell workload, slab distances carry an
    absolute rounding error of ~|o/d| 2^-24, large next to the distance to a
    nearby box; the walk folds a margin for it into the ray's slab offsets.
    Rays spawned on the surfaces:
    * the FP32 walk's closest t equals FP32 brute force's bit for bit (no box
      culled); ids may differ only at exactly equal t (the boxes' bottom
      faces lie in the floor's plane, y = 0);
    * against FP64 (reference-exact): hits on the same surface agree along
      the surface normal to 2e-3 (~32 float ulps at 555 units; at grazing
      incidence t itself moves by ~1e-2), and rays where the two precisions
      hit different surfaces (FP32 Möller–Trumbore is not watertight: a ray
      through a shared edge can slip between the triangles) stay below 32 in
      4 Mi (~12 seen);
    * FP32 any-hit below 0.999 t of the FP64 hit agrees with FP64's answer
      on all but 1e-4 of the rays."""
    o = orc.OracleScene(scenes.load("cornell_c1.json"))
    dev = ctx.scene(o.desc)
    try:
        rays = cornell_surface_rays(o, 1 << 22, seed=5)
        differ, ties = f32_walk_vs_brute(ctx, dev, rays)
        h64, _ = ctx.trace(dev, rays, mode=0, precision=rtg.PRECISION_F64)
        h32, _ = ctx.trace(dev, rays, mode=0, precision=rtg.PRECISION_F32)
        hit64, hit32 = (h64["flags"] & 1) != 0, (h32["flags"] & 1) != 0
        prims = orc.arrays(o.desc)["prims"].view(np.float64).reshape(-1, 20)
        idx = np.where(hit64, h64["prim"], 0).astype(np.int64)
        v0, v1, v2 = prims[idx, 2:5], prims[idx, 5:8], prims[idx, 8:11]
        n = np.cross(v1 - v0, v2 - v0)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        cos = np.abs((n * rays[:, 3:6]).sum(1))
        both = hit64 & hit32
        off = np.zeros(len(rays))
        off[both] = np.abs(h32["t"][both] - h64["t"][both]) * cos[both]
        other = (hit64 != hit32) | (off > 2e-3)
        print(f"cornell-scale rays: {int(hit64.sum())} fp64 hits; fp32 walk vs brute: {differ} differ, {ties} ties; "
              f"fp32 vs fp64: max distance along the normal {off[both & ~other].max():.2e}, "
              f"{int(other.sum())} on another surface", flush=True)
        assert differ == 0
        assert int(other.sum()) < 32
        r2 = rays.copy()
        r2[:, 7] = np.where(hit64, h64["t"] * 0.999, 1e4)
        occ32, _ = ctx.trace(dev, r2, mode=1, precision=rtg.PRECISION_F32)
        occ64, _ = ctx.trace(dev, r2, mode=1, precision=rtg.PRECISION_F64)
        diff = int(((occ32["flags"] & 1) != (occ64["flags"] & 1)).sum())
        print(f"cornell-scale any-hit disagreements: {diff}", flush=True)
        assert diff <= 1e-4 * len(rays)
    finally:
        dev.release()
