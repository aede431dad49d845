"""Per-kernel profiling (rtg_set_profiling / rtg_get_profile), which bench.py
reads for the closest-hit kernel's roofline. Timing every launch must not
change a render (images and counters bitwise equal, single renders and
batches alike); every launch is counted once under its kind with a positive
device time; the work totals are the calls' own counters."""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")
PASSES = [(0, 1), (1, 2), (2, 4)]


def render(ctx, o, cfg, profile, kinds=None):
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    rtg.set_profiling(ctx, profile, kinds=kinds)
    stats = [ctx.render_pass(dev, cfg, s0, s1, acc)[0] for s0, s1 in PASSES]
    prof = rtg.get_profile(ctx).as_dict()
    rtg.set_profiling(ctx, False)
    s, c = acc.read()
    acc.release()
    dev.release()
    return s, c, stats, prof


@pytest.mark.parametrize("precision", [rtg.PRECISION_F32, rtg.PRECISION_F64])
def test_profiled_render_is_unchanged_and_counted(ctx, precision):
    o = orc.OracleScene(scenes.load("torus_c2.json"), width=96, height=64)
    cfg = o.settings(precision)
    s0, c0, st0, _ = render(ctx, o, cfg, False)
    s1, c1, st1, prof = render(ctx, o, cfg, True)
    assert np.array_equal(c1, c0)
    assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64))
    for a, b in zip(st1, st0):
        for k in COUNTERS:
            assert getattr(a, k) == getattr(b, k), k
    # Each pass is one chunk here: one generate and one accumulate, and one
    # closest-hit, shade and any-hit launch per bounce.
    n = len(PASSES)
    assert prof["generate"]["launches"] == n and prof["accumulate"]["launches"] == n
    for k in ("extend", "shade", "shadow"):
        assert prof[k]["launches"] == n * cfg.max_depth, k
    for k in ("generate", "extend", "shade", "shadow", "accumulate"):
        assert prof[k]["ms"] > 0.0, k
    for k in ("direct", "trace", "plan"):
        assert prof[k]["launches"] == 0 and prof[k]["ms"] == 0.0, k
    assert prof["closest_rays"] == sum(a.rays for a in st1)
    assert prof["shadow_rays"] == sum(a.shadow_rays for a in st1)
    assert prof["closest_visits"] + prof["shadow_visits"] == sum(a.node_visits for a in st1)
    assert prof["closest_tests"] + prof["shadow_tests"] == sum(a.prim_tests for a in st1)


def test_profiled_batch_is_unchanged_and_counted(ctx):
    """Batches replay captured graphs unprofiled and launch directly while
    profiling: both give the same images and counters."""
    specs = [(orc.OracleScene(scenes.load(n), width=w, height=h), k) for n, w, h, k in
             (("torus_c2.json", 64, 48, 3), ("cornell_c1.json", 40, 40, 4))]
    out = {}
    for profile in (False, True):
        devs = [ctx.scene(o.desc) for o, _ in specs]
        cfgs = [o.settings(rtg.PRECISION_F32) for o, _ in specs]
        accs = [ctx.accum(c.width, c.height) for c in cfgs]
        rtg.set_profiling(ctx, profile)
        stats = ctx.render_batch([(d, a, c, 0, k) for d, a, c, (_, k) in zip(devs, accs, cfgs, specs)])
        prof = rtg.get_profile(ctx).as_dict()
        rtg.set_profiling(ctx, False)
        out[profile] = ([a.read() for a in accs], stats, prof)
        for x in accs + devs:
            x.release()
    (r1, st1, prof), (r0, st0, _) = out[True], out[False]
    for (s1, c1), (s0, c0) in zip(r1, r0):
        assert np.array_equal(c1, c0)
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64))
    for a, b in zip(st1, st0):
        for k in COUNTERS:
            assert getattr(a, k) == getattr(b, k), k
    depth = max(o.settings(rtg.PRECISION_F32).max_depth for o, _ in specs)
    # One batched pass: one launch per bounce of each traversal kernel.
    for k in ("extend", "shade", "shadow"):
        assert prof[k]["launches"] == depth and prof[k]["ms"] > 0.0, k
    assert prof["plan"]["launches"] == depth + 1
    assert prof["closest_rays"] == sum(a.rays for a in st1)
    assert prof["shadow_rays"] == sum(a.shadow_rays for a in st1)


def test_profiling_selected_kinds(ctx):
    """rtg_set_profiling_kinds: only the chosen kernels' launches are timed
    (bench.py times the closest-hit kernel alone); the work totals are kept
    and the render is unchanged."""
    o = orc.OracleScene(scenes.load("torus_c2.json"), width=96, height=64)
    cfg = o.settings(rtg.PRECISION_F32)
    s0, c0, st0, _ = render(ctx, o, cfg, False)
    s1, c1, st1, prof = render(ctx, o, cfg, True, kinds=["extend"])
    assert np.array_equal(c1, c0)
    assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64))
    assert prof["extend"]["launches"] == len(PASSES) * cfg.max_depth and prof["extend"]["ms"] > 0.0
    for k in ("generate", "shade", "shadow", "accumulate", "direct", "trace", "plan"):
        assert prof[k]["launches"] == 0 and prof[k]["ms"] == 0.0, k
    assert prof["closest_rays"] == sum(a.rays for a in st1)
    assert prof["shadow_rays"] == sum(a.shadow_rays for a in st1)
    with pytest.raises(rtg.RtgError, match="unknown kernel kind"):
        ctx._check(ctx.lib.rtg_set_profiling_kinds(ctx.h, 1 << len(rtg.K_NAMES)))
