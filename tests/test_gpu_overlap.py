"""Concurrency inside a pass: the any-hit kernels run on a side stream beside
the following bounces' kernels (single renders and batched passes), and an
FP32 pass runs as chunks alternating between two lanes (stream pairs with
their own path buffers). Every path's additions keep their order and every
pixel's chunks are added in sample order, so renders must be bitwise
identical with and without either (RTG_OVERLAP=0, RTG_LANES=1), counters
included, also when a pass is cut into many small chunks."""
import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")


def render(ctx, o, cfg, passes):
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    stats = [ctx.render_pass(dev, cfg, s0, s1, acc)[0] for s0, s1 in passes]
    s, c = acc.read()
    acc.release()
    dev.release()
    return s, c, stats


@pytest.mark.parametrize("name,w,h,precision", [("torus_c2.json", 160, 120, rtg.PRECISION_F32),
                                                ("sphere_field_c3.json", 128, 72, rtg.PRECISION_F32),
                                                ("cornell_c1.json", 96, 96, rtg.PRECISION_F32),
                                                ("torus_c2.json", 64, 48, rtg.PRECISION_F64),
                                                ("sphere_field_c3.json", 96, 54, "direct")])
def test_single_render_overlap_is_invisible(ctx, monkeypatch, name, w, h, precision):
    if precision == "direct":  # the direct integrator (one kernel per chunk)
        o = orc.OracleScene(scenes.load(name), width=w, height=h, integrator="direct")
        precision = rtg.PRECISION_F32
    else:
        o = orc.OracleScene(scenes.load(name), width=w, height=h)
    cfg = o.settings(precision)
    passes = [(0, 1), (1, 2), (2, 4), (4, 8)]
    monkeypatch.setenv("RTG_OVERLAP", "0")
    monkeypatch.setenv("RTG_LANES", "1")
    s0, c0, st0 = render(ctx, o, cfg, passes)
    assert sum(a.shadow_rays for a in st0) > 0
    for overlap, lanes, chunk in (("1", "1", None), ("0", "2", "4096"), ("1", "2", "4096"), ("1", "2", "1024")):
        monkeypatch.setenv("RTG_OVERLAP", overlap)
        monkeypatch.setenv("RTG_LANES", lanes)
        if chunk:
            monkeypatch.setenv("RTG_CHUNK_PATHS", chunk)
        s1, c1, st1 = render(ctx, o, cfg, passes)
        case = (overlap, lanes, chunk)
        assert np.array_equal(c1, c0), case
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64)), case
        for a, b in zip(st1, st0):
            for k in COUNTERS:
                assert getattr(a, k) == getattr(b, k), (case, k)


def test_batched_overlap_is_invisible(ctx, monkeypatch):
    specs = [(orc.OracleScene(scenes.load(n), width=w, height=h), k) for n, w, h, k in
             (("torus_c2.json", 64, 48, 3), ("cornell_c1.json", 40, 40, 4), ("sphere_field_c3.json", 48, 27, 2))]
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("RTG_OVERLAP", mode)
        devs = [ctx.scene(o.desc) for o, _ in specs]
        cfgs = [o.settings(rtg.PRECISION_F32) for o, _ in specs]
        accs = [ctx.accum(c.width, c.height) for c in cfgs]
        stats = ctx.render_batch([(d, a, c, 0, k) for d, a, c, (_, k) in zip(devs, accs, cfgs, specs)])
        out[mode] = ([a.read() for a in accs], stats)
        for x in accs + devs:
            x.release()
    (r1, st1), (r0, st0) = out["1"], out["0"]
    for (s1, c1), (s0, c0) in zip(r1, r0):
        assert np.array_equal(c1, c0)
        assert np.array_equal(s1.view(np.uint64), s0.view(np.uint64))
    for a, b in zip(st1, st0):
        for k in COUNTERS:
            assert getattr(a, k) == getattr(b, k), k
