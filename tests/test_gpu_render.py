"""GPU parity of the wavefront path tracer (rtg_render_pass).

* FP64 mode consumes the reference's PCG32 stream in double: every per-pixel
  sum must equal the oracle's to 1e-12 absolute (device sin/cos may differ
  from glibc in the last bit, which moves a path's direction by ~1e-16; the
  largest difference seen is ~1.3e-15), sample counts and ray, shadow-ray
  and roulette counters exactly.
* FP32 mode (counter-hash RNG) must converge to the closed forms of
  tests/test_render.cpp:89-136 and agree with the oracle's image statistics.
"""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu


def gpu_render(ctx, o, s0, s1, precision, passes=None):
    cfg = o.settings(precision)
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    stats = []
    for a, b in (passes or [(s0, s1)]):
        st, cancelled = ctx.render_pass(dev, cfg, a, b, acc)
        assert not cancelled
        stats.append(st.as_dict())
    s, c = acc.read()
    acc.release()
    dev.release()
    return s, c, stats


def assert_f64_equal(name, sg, cg, stg, so, co, sto):
    """FP64 device render == oracle render: counts exactly, sums to 1e-12
    absolute, and the counters of the reference's counting rules exactly."""
    print(name, "max abs diff", float(np.abs(sg - so).max()), "rays", stg["rays"], sto["rays"], "shadow",
          stg["shadow_rays"], sto["shadow_rays"], flush=True)
    assert np.array_equal(cg, co)
    assert np.abs(sg - so).max() <= 1e-12
    for k in ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples"):
        assert stg[k] == sto[k], (k, stg[k], sto[k])


@pytest.mark.parametrize("name,w,h,spp", [("cornell_box.json", 24, 24, 4), ("first_light.json", 24, 16, 3),
                                          ("materials.json", 24, 12, 3), ("sphere_field.json", 24, 16, 2),
                                          ("soft_shadows.json", 24, 16, 2), ("cornell_glass_metal.json", 24, 24, 2)])
def test_f64_render_matches_oracle(ctx, name, w, h, spp):
    j = scenes.load_asset(name)
    o = orc.OracleScene(j, width=w, height=h)
    so, co, sto = o.render(0, spp)
    sg, cg, stg = gpu_render(ctx, o, 0, spp, rtg.PRECISION_F64)
    assert_f64_equal(name, sg, cg, stg[0], so, co, sto)


@pytest.mark.parametrize("name,w,h,spp", [("cornell_c1.json", 24, 24, 4), ("torus_c2.json", 24, 24, 2),
                                          ("sphere_field_c3.json", 32, 18, 2)])
def test_f64_workload_scenes_match_oracle(ctx, name, w, h, spp):
    o = orc.OracleScene(scenes.load(name), width=w, height=h)
    so, co, sto = o.render(0, spp)
    sg, cg, stg = gpu_render(ctx, o, 0, spp, rtg.PRECISION_F64)
    assert_f64_equal(name, sg, cg, stg[0], so, co, sto)


def test_f64_render_passes_accumulate_in_sample_order(ctx):
    o = orc.OracleScene(scenes.load_asset("cornell_box.json"), width=16, height=16)
    s1, c1, _ = gpu_render(ctx, o, 0, 8, rtg.PRECISION_F64)
    s2, c2, _ = gpu_render(ctx, o, 0, 8, rtg.PRECISION_F64, passes=[(0, 1), (1, 2), (2, 4), (4, 8)])
    assert np.array_equal(c1, c2)
    assert np.array_equal(s1, s2)  # bitwise: same samples, same order


def test_f32_render_deterministic(ctx):
    o = orc.OracleScene(scenes.load_asset("cornell_box.json"), width=32, height=32)
    a = gpu_render(ctx, o, 0, 8, rtg.PRECISION_F32)
    b = gpu_render(ctx, o, 0, 8, rtg.PRECISION_F32, passes=[(0, 1), (1, 2), (2, 4), (4, 8)])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    o2 = orc.OracleScene(scenes.load_asset("cornell_box.json"), width=32, height=32, seed=99)
    c = gpu_render(ctx, o2, 0, 8, rtg.PRECISION_F32)
    assert not np.array_equal(a[0], c[0])


def mean_lum(s, c):
    return float((s[..., 1] / c).mean())


@pytest.mark.parametrize("precision", [rtg.PRECISION_F32, rtg.PRECISION_F64])
@pytest.mark.parametrize("nee", [True, False])
def test_sphere_light_over_plane_converges(ctx, precision, nee):
    o = orc.OracleScene(scenes.sphere_light_over_plane(nee))
    spp = o.R["spp"]
    s, c, _ = gpu_render(ctx, o, 0, spp, precision)
    assert mean_lum(s, c) == pytest.approx(0.2, rel=0.01 if nee else 0.03)


@pytest.mark.parametrize("nee", [True, False])
def test_ball_in_two_sided_shell(ctx, nee):
    o = orc.OracleScene(scenes.ball_in_shell(nee))
    s, c, _ = gpu_render(ctx, o, 0, o.R["spp"], rtg.PRECISION_F32)
    assert mean_lum(s, c) == pytest.approx(1.2, rel=0.01)


def test_uniform_sky(ctx):
    o = orc.OracleScene(scenes.sky_over_plane())
    s, c, _ = gpu_render(ctx, o, 0, o.R["spp"], rtg.PRECISION_F32)
    assert mean_lum(s, c) == pytest.approx(0.4, rel=0.02)


def test_f32_cornell_mean_within_3_sigma_of_oracle(ctx):
    """Per-channel image mean of the FP32 GPU render vs the oracle's
    (reference-exact) render of the same scene, 3 sigma of the difference."""
    o = orc.OracleScene(scenes.load("cornell_c1.json"), width=32, height=32)
    spp = 256
    so, co, _ = o.render(0, spp)
    sg, cg, _ = gpu_render(ctx, o, 0, spp, rtg.PRECISION_F32)
    mo, mg = so / co[..., None], sg / cg[..., None]
    # sigma of the image mean from per-pixel variation across two halves
    s_a, c_a, _ = gpu_render(ctx, o, 0, spp // 2, rtg.PRECISION_F32)
    s_b, c_b, _ = gpu_render(ctx, o, spp // 2, spp, rtg.PRECISION_F32)
    half_diff = (s_a / c_a[..., None] - s_b / c_b[..., None]).reshape(-1, 3)
    # Var(half_diff) = 2 var/(spp/2) = 4 var/spp: its sd / 2 is the per-pixel
    # sd of a spp-sample mean (RMS over pixels)
    sigma_pix = half_diff.std(axis=0) / 2
    # the difference of two independent spp-sample image means
    sigma_mean = sigma_pix / np.sqrt(mo.shape[0] * mo.shape[1]) * np.sqrt(2)
    d = mg.reshape(-1, 3).mean(0) - mo.reshape(-1, 3).mean(0)
    print("mean diff", d, "sigma", sigma_mean, "rmse", np.sqrt(((mg - mo) ** 2).mean()), flush=True)
    assert (np.abs(d) <= 3 * sigma_mean + 1e-9).all()


@pytest.mark.parametrize("integrator", ["direct", "normals"])
def test_direct_and_normals_f64_match_oracle(ctx, integrator):
    j = scenes.load_asset("materials.json")
    o = orc.OracleScene(j, width=20, height=12, integrator=integrator)
    so, co, sto = o.render(0, 2)
    sg, cg, stg = gpu_render(ctx, o, 0, 2, rtg.PRECISION_F64)
    assert_f64_equal(integrator, sg, cg, stg[0], so, co, sto)


def test_no_bvh_render_matches_bvh_render(ctx):
    o = orc.OracleScene(scenes.load("cornell_c1.json"), width=16, height=16)
    a, _, _ = gpu_render(ctx, o, 0, 4, rtg.PRECISION_F64)
    o2 = orc.OracleScene(scenes.load("cornell_c1.json"), width=16, height=16, use_bvh=False)
    b, _, _ = gpu_render(ctx, o2, 0, 4, rtg.PRECISION_F64)
    # scene.cpp:200-215 finds the same closest hits as the BVH
    assert np.abs(a - b).max() <= 1e-12


@pytest.mark.parametrize("integrator", ["path", "direct"])
def test_f32_no_bvh_render_is_brute_force(ctx, integrator):
    """--no-bvh in FP32 (the persistent traversal kernels and the direct
    queue): every closest-hit ray tests every primitive (scene.cpp:200-215)
    and no node box, and the image equals the BVH render's bit for bit (the
    FP32 walk's closest t equals brute force's; the sphere field has no
    equal-t ties)."""
    kw = {} if integrator == "path" else {"integrator": "direct"}
    o = orc.OracleScene(scenes.load("sphere_field_c3.json"), width=48, height=27, **kw)
    ob = orc.OracleScene(scenes.load("sphere_field_c3.json"), width=48, height=27, use_bvh=False, **kw)
    a, ca, sta = gpu_render(ctx, o, 0, 2, rtg.PRECISION_F32)
    b, cb, stb = gpu_render(ctx, ob, 0, 2, rtg.PRECISION_F32)
    st_a, st_b = sta[0], stb[0]
    n = ob.desc.n_prims
    assert st_b["node_visits"] == 0
    assert st_b["rays"] * n <= st_b["prim_tests"] <= (st_b["rays"] + st_b["shadow_rays"]) * n
    assert (st_a["rays"], st_a["shadow_rays"]) == (st_b["rays"], st_b["shadow_rays"])
    assert np.array_equal(ca, cb) and np.array_equal(a.view(np.uint64), b.view(np.uint64))
