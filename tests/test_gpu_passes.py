"""rtg_render_passes: a progressive schedule in one call, the passes running
side by side on the device (or, with no cancel flag, merged into one item of
all their samples). Every pixel's additions must stay in sample
order, so the framebuffer (sums and counts) and the counters must be bitwise
those of rtg_render_pass called once per pass — in both precisions, for the
direct integrator (one lane), with one lane forced, and with passes cut into
many small chunks."""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")


def schedule(spp, max_pass=32):
    b = [0]
    while b[-1] < spp:
        b.append(b[-1] + min(spp - b[-1], max(1, b[-1]), max_pass))
    return b


def one_by_one(ctx, dev, cfg, bounds):
    acc = ctx.accum(cfg.width, cfg.height)
    stats = [ctx.render_pass(dev, cfg, a, b, acc)[0] for a, b in zip(bounds, bounds[1:])]
    s, c = acc.read()
    acc.release()
    return s, c, {k: sum(getattr(x, k) for x in stats) for k in COUNTERS}


def in_one_call(ctx, dev, cfg, bounds):
    acc = ctx.accum(cfg.width, cfg.height)
    st, done, cancelled = ctx.render_passes(dev, cfg, bounds, acc)
    assert (done, cancelled) == (len(bounds) - 1, False)
    s, c = acc.read()
    acc.release()
    return s, c, {k: getattr(st, k) for k in COUNTERS}


@pytest.mark.parametrize("name,w,h,spp,precision,integrator", [
    ("torus_c2.json", 128, 96, 16, rtg.PRECISION_F32, None),
    ("sphere_field_c3.json", 128, 72, 8, rtg.PRECISION_F32, None),
    ("cornell_c1.json", 96, 96, 8, rtg.PRECISION_F32, None),
    ("torus_c2.json", 64, 48, 8, rtg.PRECISION_F64, None),
    ("sphere_field_c3.json", 96, 54, 4, rtg.PRECISION_F32, "direct"),
])
def test_passes_in_one_call_equal_pass_by_pass(ctx, monkeypatch, name, w, h, spp, precision, integrator):
    kw = {"integrator": integrator} if integrator else {}
    o = orc.OracleScene(scenes.load(name), width=w, height=h, **kw)
    cfg = o.settings(precision)
    dev = ctx.scene(o.desc)
    bounds = schedule(spp)
    s0, c0, k0 = one_by_one(ctx, dev, cfg, bounds)
    assert int(c0.min()) == spp and k0["rays"] > 0
    # merged: the passes run as one item of all their samples (no cancel
    # flag); "0": each pass an item of its own
    for lanes, chunk, merge in (("2", None, "1"), ("2", None, "0"), ("1", None, "1"), ("2", "2048", "1"),
                                ("2", "2048", "0")):
        monkeypatch.setenv("RTG_LANES", lanes)
        monkeypatch.setenv("RTG_MERGE_PASSES", merge)
        if chunk:
            monkeypatch.setenv("RTG_CHUNK_PATHS", chunk)
        s1, c1, k1 = in_one_call(ctx, dev, cfg, bounds)
        case = (lanes, chunk, merge)
        assert np.array_equal(c1, c0), case
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64)), case
        assert k1 == k0, case
    dev.release()


def test_passes_cancel_and_bad_bounds(ctx):
    o = orc.OracleScene(scenes.load("cornell_c1.json"), width=48, height=48)
    cfg = o.settings(rtg.PRECISION_F32)
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    flag = rtg.C.c_int32(1)
    st, done, cancelled = ctx.render_passes(dev, cfg, schedule(8), acc, cancel=flag)
    assert cancelled and done == 0 and st.rays == 0
    assert int(acc.read()[1].max()) == 0
    flag.value = 0
    st, done, cancelled = ctx.render_passes(dev, cfg, schedule(8), acc, cancel=flag)
    assert not cancelled and done == len(schedule(8)) - 1 and int(acc.read()[1].min()) == 8
    for bad in ([0, 2, 2], [0, 4, 3]):
        with pytest.raises(rtg.RtgError, match="empty sample range"):
            ctx.render_passes(dev, cfg, bad, acc)
    acc.release()
    dev.release()


@pytest.mark.parametrize("name,w,h,spp", [("torus_c2.json", 128, 96, 8), ("sphere_field_c3.json", 128, 72, 4),
                                         ("cornell_c1.json", 96, 96, 8)])
def test_persistent_traversal_equals_one_ray_per_thread(ctx, monkeypatch, name, w, h, spp):
    """The persistent traversal kernels hand idle lanes new rays between
    rounds of their walks; a ray's walk depends on the ray alone, so the
    framebuffer and every counter equal those of one ray per thread
    (RTG_PERSISTENT=0) bit for bit, also with chunks cut small."""
    o = orc.OracleScene(scenes.load(name), width=w, height=h)
    cfg = o.settings(rtg.PRECISION_F32)
    dev = ctx.scene(o.desc)
    bounds = schedule(spp)
    monkeypatch.setenv("RTG_PERSISTENT", "0")
    s0, c0, k0 = in_one_call(ctx, dev, cfg, bounds)
    for env in ({"RTG_PERSISTENT": "1"}, {"RTG_PERSISTENT": "1", "RTG_CHUNK_PATHS": "4096"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        s1, c1, k1 = in_one_call(ctx, dev, cfg, bounds)
        assert np.array_equal(c1, c0), env
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64)), env
        assert k1 == k0, env
    dev.release()
