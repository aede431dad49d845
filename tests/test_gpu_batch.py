"""rtg_render_batch: many jobs (scenes, resolutions, sample ranges, depths)
in shared launches. A job's samples are keyed on (seed, pixel, sample,
bounce) and its sums are added in sample order, so every batched job must
equal the same pass rendered alone with rtg_render_pass bit for bit, and its
counters (rays, shadow rays, roulette, node visits, primitive tests) must be
identical."""
import os

import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")


def job_spec(j, w, h, spp, seed=None, depth=None, precision=rtg.PRECISION_F32, integrator=None):
    over = {"width": w, "height": h}
    if seed is not None:
        over["seed"] = seed
    if integrator is not None:
        over["integrator"] = integrator
    o = orc.OracleScene(j, **over)
    cfg = o.settings(precision)
    if depth is not None:
        cfg.max_depth = depth
    return o, cfg, spp


@pytest.fixture(scope="module")
def specs():
    return [
        job_spec(scenes.load("cornell_c1.json"), 32, 32, 4),
        job_spec(scenes.load("torus_c2.json"), 24, 20, 3),
        job_spec(scenes.load("sphere_field_c3.json"), 40, 24, 2),
        job_spec(scenes.load_asset("materials.json"), 24, 12, 5),
        job_spec(scenes.load_asset("cornell_box.json"), 17, 13, 7, seed=5, depth=3),
        job_spec(scenes.load_asset("first_light.json"), 16, 16, 1),
        job_spec(scenes.load_asset("cornell_glass_metal.json"), 20, 20, 2, depth=12),
    ]


def solo(ctx, o, cfg, s0, s1, p=None):
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    st, _ = ctx.render_pass(dev, cfg, s0, s1, acc)
    s, c = acc.read()
    acc.release()
    dev.release()
    return s, c, st


def batched(ctx, specs, ranges):
    devs = [ctx.scene(o.desc) for o, _, _ in specs]
    accs = [ctx.accum(cfg.width, cfg.height) for _, cfg, _ in specs]
    out = []
    for s0, s1 in ranges:
        jobs = [(d, a, cfg, s0, min(s1, spp)) for d, a, (o, cfg, spp) in zip(devs, accs, specs) if s0 < spp]
        out.append(ctx.render_batch(jobs))
    res = [a.read() for a in accs]
    for a in accs:
        a.release()
    for d in devs:
        d.release()
    return res, out


def test_batch_equals_solo_renders(ctx, specs):
    (res, (stats,)) = batched(ctx, specs, [(0, 1000)])
    for (o, cfg, spp), (s, c), st in zip(specs, res, stats):
        s1, c1, st1 = solo(ctx, o, cfg, 0, spp)
        assert np.array_equal(c, c1)
        assert np.array_equal(s.view(np.uint64), s1.view(np.uint64))
        for k in COUNTERS:
            assert getattr(st, k) == getattr(st1, k), k


def test_progressive_batches_equal_one_solo_pass(ctx, specs):
    """Passes 1, 1, 2, ... of every job batched together = one solo pass."""
    res, _ = batched(ctx, specs, [(0, 1), (1, 2), (2, 4), (4, 8)])
    for (o, cfg, spp), (s, c) in zip(specs, res):
        s1, c1, _ = solo(ctx, o, cfg, 0, spp)
        assert np.array_equal(c, c1) and np.array_equal(s, s1)


def test_pixel_ranges_of_one_buffer(ctx):
    o, cfg, _ = job_spec(scenes.load("cornell_c1.json"), 30, 20, 3)
    dev = ctx.scene(o.desc)
    acc = ctx.accum(30, 20)
    st = ctx.render_batch([(dev, acc, cfg, 0, 3, 0, 7), (dev, acc, cfg, 0, 3, 7, 450), (dev, acc, cfg, 0, 3, 450, 600)])
    s, c = acc.read()
    with pytest.raises(rtg.RtgError, match="overlapping"):
        ctx.render_batch([(dev, acc, cfg, 3, 4, 0, 10), (dev, acc, cfg, 3, 4, 9, 20)])
    acc.release()
    dev.release()
    s1, c1, st1 = solo(ctx, o, cfg, 0, 3)
    assert np.array_equal(c, c1) and np.array_equal(s, s1)
    assert sum(x.rays for x in st) == st1.rays


def test_jobs_outside_the_batch_kernels(ctx):
    """FP64 and direct/normals jobs go through the single-render path; a job
    of 4 Mi paths shares a wide batch pass (grids over every path) with the
    small one."""
    sp = [job_spec(scenes.load_asset("materials.json"), 20, 12, 2, precision=rtg.PRECISION_F64),
          job_spec(scenes.load_asset("materials.json"), 20, 12, 2, integrator="direct"),
          job_spec(scenes.load_asset("mirrors.json"), 16, 8, 2, integrator="normals"),
          job_spec(scenes.load("cornell_c1.json"), 1024, 1024, 4),
          job_spec(scenes.load("cornell_c1.json"), 16, 16, 2)]
    res, (stats,) = batched(ctx, sp, [(0, 4)])
    for (o, cfg, spp), (s, c), st in zip(sp, res, stats):
        s1, c1, st1 = solo(ctx, o, cfg, 0, spp)
        assert np.array_equal(c, c1) and np.array_equal(s, s1)
        # the single-render jobs run side by side, each with its own counters
        for k in COUNTERS:
            assert getattr(st, k) == getattr(st1, k), k


def test_many_small_jobs_in_several_groups(ctx, specs, monkeypatch):
    """300 jobs (more than one group holds) and a small capacity that splits
    jobs into segments of whole pixels and sample sub-ranges."""
    monkeypatch.setenv("RTG_CHUNK_PATHS", "4096")
    pool = [specs[i][0].json for i in (0, 2, 3, 4, 5, 6)]  # the small scenes
    many = [job_spec(pool[i % len(pool)], 8 + i % 5, 6 + i % 3, 1 + i % 3, seed=i) for i in range(300)]
    many[7] = job_spec(scenes.load("cornell_c1.json"), 64, 48, 3)  # 9216 paths: split into segments
    res, (stats,) = batched(ctx, many, [(0, 3)])
    for i in (0, 1, 7, 123, 299):
        o, cfg, spp = many[i]
        s1, c1, st1 = solo(ctx, o, cfg, 0, spp)
        assert np.array_equal(res[i][1], c1) and np.array_equal(res[i][0], s1), i
        assert stats[i].rays == st1.rays and stats[i].node_visits == st1.node_visits


def test_sample_ranges_of_one_pixel_across_lanes(ctx, monkeypatch):
    """Chunks of 1024 paths and 2500 samples per pixel: each pixel's samples
    are cut into sub-ranges that land in groups on both lanes; the pixel's
    additions must still come in sample order (bitwise the solo render)."""
    monkeypatch.setenv("RTG_CHUNK_PATHS", "1024")
    sp = [job_spec(scenes.load("cornell_c1.json"), 3, 1, 2500), job_spec(scenes.load_asset("first_light.json"), 2, 2, 700)]
    res, (stats,) = batched(ctx, sp, [(0, 2500)])
    for (o, cfg, spp), (s, c), st in zip(sp, res, stats):
        s1, c1, st1 = solo(ctx, o, cfg, 0, spp)
        assert int(c.min()) == spp
        assert np.array_equal(c, c1) and np.array_equal(s.view(np.uint64), s1.view(np.uint64))
        for k in COUNTERS:
            assert getattr(st, k) == getattr(st1, k), k


def test_graph_replays_equal_direct_launches(ctx, specs, monkeypatch):
    """A batch pass replays as a captured graph (one per max depth and stack
    class, reused across calls). Replays in a row, with eager renders in
    between, must each equal the same batch launched kernel by kernel
    (RTG_BATCH_GRAPHS=0), and a different job mix must reuse the graph."""
    for rnd, mix in enumerate((specs, specs[:3], specs, specs[2:])):
        monkeypatch.setenv("RTG_BATCH_GRAPHS", "1")
        res_g, (st_g,) = batched(ctx, mix, [(rnd, rnd + 2)])
        o, cfg, spp = mix[0]
        solo(ctx, o, cfg, 0, 1)  # eager work between the replays
        monkeypatch.setenv("RTG_BATCH_GRAPHS", "0")
        res_d, (st_d,) = batched(ctx, mix, [(rnd, rnd + 2)])
        for i, ((sg, cg), (sd, cd)) in enumerate(zip(res_g, res_d)):
            assert np.array_equal(cg, cd), (rnd, i)
            assert np.array_equal(sg.view(np.uint64), sd.view(np.uint64)), (rnd, i)
        assert len(st_g) == len(st_d) and any(c.any() for _, c in res_g)
        for i, (a, b) in enumerate(zip(st_g, st_d)):
            for k in COUNTERS:
                assert getattr(a, k) == getattr(b, k), (rnd, i, k)


def test_autotune_keeps_images_and_caches_per_device(ctx, specs, tmp_path):
    """The autotuner times candidate knobs on a workload of batch calls; the
    images do not depend on the knobs it picks, and a second call with the
    same cache file and key reuses the result without timing."""
    devs = [ctx.scene(o.desc) for o, _, _ in specs]
    accs = [ctx.accum(cfg.width, cfg.height) for _, cfg, _ in specs]
    calls = [[(d, a, cfg, 0, 1) for d, a, (o, cfg, spp) in zip(devs, accs, specs)],
             [(d, a, cfg, 1, spp) for d, a, (o, cfg, spp) in zip(devs, accs, specs) if spp > 1]]
    ctx.set_tuning()
    for call in calls:
        ctx.render_batch(call)
    ref = [a.read() for a in accs]
    cache = str(tmp_path / "tune.jsonl")
    r1 = ctx.autotune(calls, cache_path=cache, key="small-mix")
    print("autotune", r1, flush=True)
    assert r1["trials"] > 0 and r1["best_ms"] > 0 and r1["default_ms"] > 0
    assert ctx.get_tuning() == r1["tuning"]
    assert all(int(a.read()[1].max()) == 0 for a in accs)  # left cleared
    for call in calls:
        ctx.render_batch(call)
    for (s, c), (s0, c0) in zip((a.read() for a in accs), ref):
        assert np.array_equal(c, c0) and np.array_equal(s.view(np.uint64), s0.view(np.uint64))
    ctx.set_tuning()
    r2 = ctx.autotune(calls, cache_path=cache, key="small-mix")
    assert r2["trials"] == 0 and r2["tuning"] == r1["tuning"] and ctx.get_tuning() == r1["tuning"]
    with pytest.raises(rtg.RtgError, match="key"):
        ctx.autotune(calls, cache_path=cache, key='bad"key')
    ctx.set_tuning()
    for x in accs + devs:
        x.release()


def test_bad_jobs_are_rejected(ctx, specs):
    o, cfg, _ = specs[0]
    dev = ctx.scene(o.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    with pytest.raises(rtg.RtgError, match="empty sample range"):
        ctx.render_batch([(dev, acc, cfg, 2, 2)])
    with pytest.raises(rtg.RtgError, match="bad pixel range"):
        ctx.render_batch([(dev, acc, cfg, 0, 1, 5, cfg.width * cfg.height + 1)])
    small = ctx.accum(4, 4)
    with pytest.raises(rtg.RtgError, match="size differs"):
        ctx.render_batch([(dev, small, cfg, 0, 1)])
    small.release()
    acc.release()
    dev.release()
