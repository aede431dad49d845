"""rtg_render_pass_static, the render pass that allocates no device memory
(path state in the library's fixed device scratch), against rtg_render_pass
on the same scene, settings and samples: framebuffer sums, counts and every
counter must be bitwise equal. Two calls in a row share the scratch, the
second on a smaller input, so state the first call left behind must not leak
into the second; a pass larger than the scratch's chunk runs as several
chunks."""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")


def render(ctx, dev, cfg, s0, s1, static, stats=True):
    acc = ctx.accum(cfg.width, cfg.height)
    if static:
        st, cancelled = ctx.render_pass_static(dev, cfg, s0, s1, acc, stats=stats)
    else:
        st, cancelled = ctx.render_pass(dev, cfg, s0, s1, acc)
    assert not cancelled
    s, c = acc.read()
    acc.release()
    return s, c, ({k: getattr(st, k) for k in COUNTERS} if st is not None else None)


# (scene, width, height, s0, s1, precision, integrator): the first case is
# 614,400 FP32 paths, more than one 256 Ki-path chunk of the scratch; each
# case after it is smaller than the one before.
CASES = [
    ("torus_c2.json", 640, 480, 0, 2, rtg.PRECISION_F32, None),
    ("cornell_c1.json", 96, 64, 1, 5, rtg.PRECISION_F32, None),
    ("torus_c2.json", 320, 240, 0, 4, rtg.PRECISION_F64, None),
    ("sphere_field_c3.json", 64, 36, 0, 3, rtg.PRECISION_F32, "direct"),
]


def test_static_pass_equals_render_pass_on_consecutive_calls(ctx):
    for name, w, h, s0, s1, precision, integrator in CASES:
        print("case", name, w, h, s0, s1, precision, integrator, flush=True)
        kw = {"integrator": integrator} if integrator else {}
        o = orc.OracleScene(scenes.load(name), width=w, height=h, **kw)
        cfg = o.settings(precision)
        dev = ctx.scene(o.desc)
        s_ref, c_ref, k_ref = render(ctx, dev, cfg, s0, s1, static=False)
        s_st, c_st, k_st = render(ctx, dev, cfg, s0, s1, static=True)
        assert int(c_ref.min()) == s1 - s0 and k_ref["rays"] > 0
        assert np.array_equal(c_st, c_ref)
        assert np.array_equal(s_st.view(np.uint64), s_ref.view(np.uint64))
        assert k_st == k_ref
        dev.release()


def test_static_pass_enqueue_only_and_repeat(ctx):
    """stats=None, cancel=None: the call only enqueues; reading the buffer
    orders after it. A repeat of the same pass gives the same bits again."""
    o = orc.OracleScene(scenes.load("torus_c2.json"), width=160, height=120)
    cfg = o.settings(rtg.PRECISION_F32)
    dev = ctx.scene(o.desc)
    s_ref, c_ref, _ = render(ctx, dev, cfg, 0, 8, static=False)
    for _ in range(2):
        s_st, c_st, k = render(ctx, dev, cfg, 0, 8, static=True, stats=False)
        assert k is None
        assert np.array_equal(c_st, c_ref)
        assert np.array_equal(s_st.view(np.uint64), s_ref.view(np.uint64))
    dev.release()
