"""The direct integrator far past 31 bounces (the reference allows max_depth
up to 1000, src/io/scene_io.cpp validation; integrator.cpp:141-174 recurses
once per metal bounce and twice per dielectric hit). FP64 device renders
must equal the reference's own renders (tests/golden/reference_direct_deep.npz,
written by make_golden.py --direct-deep from the reference build): a corridor
of two facing mirrors at depth 1000 and 200, glass and metal at depth 40.
FP32 renders complete every sample with finite sums."""
import json
import os

import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DEEP = np.load(os.path.join(HERE, "golden", "reference_direct_deep.npz"))
CASES = [("mirror_corridor.json", 12, 8, 2, 1000), ("mirror_corridor.json", 12, 8, 2, 200),
         ("cornell_glass_metal.json", 10, 10, 2, 40)]


def scene(name):
    if name == "mirror_corridor.json":
        with open(os.path.join(HERE, "golden", name)) as f:
            return json.load(f)
    return scenes.load_asset(name)


@pytest.mark.parametrize("name,w,h,spp,depth", CASES)
def test_deep_direct_fp64_equals_reference(ctx, name, w, h, spp, depth):
    key = f"{name[:-5]}_{depth}"
    o = orc.OracleScene(scene(name), width=w, height=h, integrator="direct", max_depth=depth)
    cfg = o.settings(rtg.PRECISION_F64)
    assert cfg.max_depth == depth
    dev = ctx.scene(o.desc)
    acc = ctx.accum(w, h)
    st, _ = ctx.render_pass(dev, cfg, 0, spp, acc)
    s, c = acc.read()
    want = DEEP[key + "_stats"]
    diff = float(np.abs(s - DEEP[key + "_sum"]).max())
    print(key, "max abs diff", diff, "rays", st.rays, int(want[1]), flush=True)
    assert np.array_equal(c, DEEP[key + "_count"]) and diff <= 1e-12
    assert (st.camera_samples, st.rays, st.shadow_rays) == tuple(int(x) for x in want[:3])
    # FP32: every sample completes, sums finite and close to FP64's scale
    cfg32 = o.settings(rtg.PRECISION_F32)
    acc.clear()
    st32, _ = ctx.render_pass(dev, cfg32, 0, spp, acc)
    s32, c32 = acc.read()
    assert int(c32.min()) == spp and np.isfinite(s32).all() and st32.rays > st32.camera_samples
    acc.release()
    dev.release()


def test_deep_direct_without_device_allocation(ctx):
    """rtg_render_pass_static at depth 1000: its chunk shrinks so the frames
    fit the scratch array; the result equals rtg_render_pass bit for bit."""
    o = orc.OracleScene(scene("mirror_corridor.json"), width=12, height=8, integrator="direct", max_depth=1000)
    cfg = o.settings(rtg.PRECISION_F64)
    dev = ctx.scene(o.desc)
    a1, a2 = ctx.accum(12, 8), ctx.accum(12, 8)
    st1, _ = ctx.render_pass(dev, cfg, 0, 2, a1)
    st2, _ = ctx.render_pass_static(dev, cfg, 0, 2, a2)
    (s1, c1), (s2, c2) = a1.read(), a2.read()
    assert np.array_equal(c1, c2) and np.array_equal(s1.view(np.uint64), s2.view(np.uint64))
    assert (st1.rays, st1.shadow_rays, st1.node_visits) == (st2.rays, st2.shadow_rays, st2.node_visits)
    for x in (a1, a2, dev):
        x.release()


COUNTERS = ("camera_samples", "rays", "shadow_rays", "invalid_samples", "node_visits", "prim_tests")
QUEUE_CASES = [("sphere_field_c3.json", None, 96, 54, 4, 12), ("cornell_glass_metal.json", "asset", 40, 40, 2, 40),
               ("cornell_c1.json", None, 64, 64, 2, 8)]


def load_case(name, kind):
    return scenes.load_asset(name) if kind == "asset" else scenes.load(name)


@pytest.mark.parametrize("name,kind,w,h,spp,depth", QUEUE_CASES)
def test_direct_queue_equals_depth_first_and_overflow(ctx, monkeypatch, name, kind, w, h, spp, depth):
    """FP32 direct runs as a weighted ray queue (generations of weighted
    rays, order-free fixed-point sums). Forced depth first (RTG_DIRECT=dfs),
    and with queues so small that generations overflow and their chunks are
    recomputed depth first (RTG_DIRECT_GROWTH), the framebuffer and every
    counter must be bitwise the same; chunks cut small too."""
    o = orc.OracleScene(load_case(name, kind), width=w, height=h, integrator="direct", max_depth=depth)
    cfg = o.settings(rtg.PRECISION_F32)
    dev = ctx.scene(o.desc)

    def run():
        acc = ctx.accum(w, h)
        st, _ = ctx.render_pass(dev, cfg, 0, spp, acc)
        s, c = acc.read()
        acc.release()
        return s, c, st

    s0, c0, st0 = run()
    assert int(c0.min()) == spp and st0.redone_chunks == 0 and st0.invalid_samples == 0  # default queues suffice
    assert st0.rays >= st0.camera_samples and st0.shadow_rays > 0 and np.isfinite(s0).all()
    for env in ({"RTG_DIRECT": "dfs"}, {"RTG_DIRECT_GROWTH": "0.0001"},
                {"RTG_DIRECT_GROWTH": "0.0001", "RTG_CHUNK_PATHS": "2048"}, {"RTG_CHUNK_PATHS": "2048"}):
        for k in ("RTG_DIRECT", "RTG_DIRECT_GROWTH", "RTG_CHUNK_PATHS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        s1, c1, st1 = run()
        print(name, env, "redone", st1.redone_chunks, "rays", st1.rays, st0.rays, flush=True)
        assert np.array_equal(c1, c0), env
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64)), env
        assert all(getattr(st1, k) == getattr(st0, k) for k in COUNTERS), env
        if "RTG_DIRECT_GROWTH" in env and name != "cornell_c1.json":
            assert st1.redone_chunks > 0, env  # the overflow path really ran
        elif "RTG_DIRECT" in env:
            assert st1.redone_chunks == 0, env  # forced runs count no overflow
    dev.release()


def test_direct_queue_without_device_allocation(ctx):
    """rtg_render_pass_static keeps the FP32 direct queues in the scratch
    array: bitwise the result of rtg_render_pass."""
    o = orc.OracleScene(scenes.load("sphere_field_c3.json"), width=64, height=36, integrator="direct")
    cfg = o.settings(rtg.PRECISION_F32)
    dev = ctx.scene(o.desc)
    a1, a2 = ctx.accum(64, 36), ctx.accum(64, 36)
    st1, _ = ctx.render_pass(dev, cfg, 0, 4, a1)
    st2, _ = ctx.render_pass_static(dev, cfg, 0, 4, a2)
    (s1, c1), (s2, c2) = a1.read(), a2.read()
    assert np.array_equal(c1, c2) and np.array_equal(s1.view(np.uint64), s2.view(np.uint64))
    assert all(getattr(st1, k) == getattr(st2, k) for k in COUNTERS)
    for x in (a1, a2, dev):
        x.release()
