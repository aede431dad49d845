"""The batched (C4) and served (C5) workloads at the sizes the benchmark
runs them.

C4: the 64 renders of make_batch.py seed 4 (tori of 5k-200k triangles,
Cornell variants, sphere fields; 64^2-1024^2; 1-256 spp), each with the
reference's progressive pass schedule. One rtg_render_batch per pass must
give every render bitwise the framebuffer and counters of its passes
rendered one by one with rtg_render_pass; and a few of the scenes, shrunk,
rendered in FP64 inside a batch must equal the oracle.

C5: 16 resident scenes of make_batch.py seed 5, previews of 128^2 (1-4 spp,
depth 4) and finals of 512^2-1024^2 (256 spp, progressive) arriving within
a fifth of a second. The served loop, the plain loop and each request alone
must give every request the same image hash."""
import json
import os
import sys

import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
sys.path.insert(0, scenes.SCENE_DIR)
import make_batch  # noqa: E402

COUNTERS = ("camera_samples", "rays", "shadow_rays", "rr_terminated", "invalid_samples", "node_visits", "prim_tests")


def schedule(spp, max_pass=32):
    out, done = [], 0
    while done < spp:
        k = min(spp - done, max(1, done), max_pass)
        out.append((done, done + k))
        done += k
    return out


def test_c4_batched_equals_per_render_passes(ctx):
    specs = make_batch.build(4, 64)
    hs = [rtg.HostScene(json.dumps(j)) for j in specs]
    devs = [ctx.scene(h.desc) for h in hs]
    cfgs = [h.settings(rtg.PRECISION_F32) for h in hs]
    scheds = [schedule(int(h.info()["spp"])) for h in hs]
    paths = sum(c.width * c.height * s[-1][1] for c, s in zip(cfgs, scheds))
    assert len(specs) == 64 and paths > 1_000_000_000
    # per render, pass by pass
    ref, ref_k = [], []
    for d, c, sched in zip(devs, cfgs, scheds):
        acc = ctx.accum(c.width, c.height)
        tot = dict.fromkeys(COUNTERS, 0)
        for s0, s1 in sched:
            st, _ = ctx.render_pass(d, c, s0, s1, acc)
            for k in COUNTERS:
                tot[k] += getattr(st, k)
        ref.append(acc.read())
        ref_k.append(tot)
        acc.release()
    # pass p of every unfinished render in one batch call
    accs = [ctx.accum(c.width, c.height) for c in cfgs]
    got_k = [dict.fromkeys(COUNTERS, 0) for _ in specs]
    for p in range(max(len(s) for s in scheds)):
        idx = [i for i in range(64) if p < len(scheds[i])]
        stats = ctx.render_batch([(devs[i], accs[i], cfgs[i]) + scheds[i][p] for i in idx])
        for i, st in zip(idx, stats):
            for k in COUNTERS:
                got_k[i][k] += getattr(st, k)
    for i in range(64):
        s, c = accs[i].read()
        assert np.array_equal(c, ref[i][1]), i
        assert np.array_equal(s.view(np.uint64), ref[i][0].view(np.uint64)), i
        assert got_k[i] == ref_k[i], i
    print(f"64 renders, {paths} paths, {sum(k['rays'] + k['shadow_rays'] for k in ref_k)} rays: identical",
          flush=True)
    for x in accs + devs:
        x.release()


def test_c4_scenes_in_fp64_batches_equal_the_oracle(ctx):
    """One scene of each kind from the C4 list, shrunk to 24x16 at 2 spp,
    rendered in FP64 inside one batch (with an FP32 job beside them)."""
    specs = make_batch.build(4, 64)
    picks = {}
    for j in specs:
        picks.setdefault(j["name"].split()[0], j)
    assert sorted(picks) == ["Cornell", "Sphere", "Torus"]
    jobs, oracles, accs, devs = [], [], [], []
    for kind, j in sorted(picks.items()):
        j = json.loads(json.dumps(j))
        j["render"].update(width=24, height=16, spp=2)
        hs = rtg.HostScene(json.dumps(j))
        dev = ctx.scene(hs.desc)
        acc = ctx.accum(24, 16)
        jobs.append((dev, acc, hs.settings(rtg.PRECISION_F64), 0, 2))
        oracles.append(orc.OracleScene(j, width=24, height=16))
        accs.append(acc)
        devs.append(dev)
    extra = ctx.accum(24, 16)
    jobs.append((devs[0], extra, jobs[0][2].__class__.from_buffer_copy(jobs[0][2]), 0, 2))
    jobs[-1][2].precision = rtg.PRECISION_F32
    stats = ctx.render_batch(jobs)
    for (kind, _), o, acc, st in zip(sorted(picks.items()), oracles, accs, stats):
        so, co, sto = o.render(0, 2)
        s, c = acc.read()
        diff = float(np.abs(s - so).max())
        print(kind, "max abs diff", diff, "rays", st.rays, sto["rays"], flush=True)
        assert np.array_equal(c, co) and diff <= 1e-12
        assert (st.rays, st.shadow_rays) == (sto["rays"], sto["shadow_rays"])
    for x in accs + devs + [extra]:
        x.release()


def c5_trace(seed=5, n_prev=20, finals=(512, 768, 1024, 512)):
    rng = np.random.default_rng(seed)
    specs = make_batch.build(seed, 16)
    depth = np.array([int(j["render"]["max_depth"]) for j in specs])
    n = n_prev + len(finals)
    r = np.zeros(n, rtg.REQUEST_DTYPE)
    kinds = np.zeros(n, bool)
    kinds[rng.choice(np.arange(1, n), len(finals), replace=False)] = True
    r["final_kind"] = kinds
    r["scene"] = rng.integers(0, 16, n)
    r["seed"] = rng.integers(1, 10 ** 9, n)
    size = np.full(n, 128)
    size[kinds] = finals
    r["width"] = r["height"] = size
    r["spp"] = np.where(kinds, 256, rng.integers(1, 5, n))
    r["max_depth"] = np.where(kinds, depth[r["scene"]], 4)
    r["max_pass"] = np.where(kinds, 32, 4)
    r["arrival_ms"] = np.sort(rng.uniform(0.0, 200.0, n))
    return specs, r


def test_c5_served_plain_and_alone_hash_the_same():
    specs, reqs = c5_trace()
    hs = [rtg.HostScene(json.dumps(j)) for j in specs]
    srv = rtg.Server(0, 0, 0)
    for h in hs:
        srv.add_scene(h)
    a, sa = srv.run(reqs, served=True)
    b, sb = srv.run(reqs, served=False)
    assert (a["spp_done"] == reqs["spp"]).all() and (b["spp_done"] == reqs["spp"]).all()
    assert np.array_equal(a["image_hash"], b["image_hash"])
    assert np.array_equal(a["rays"], b["rays"]) and np.array_equal(a["shadow_rays"], b["shadow_rays"])
    assert sa["max_active"] > 1  # requests overlapped on the device
    for i in range(len(reqs)):
        alone, _ = srv.run(reqs[i:i + 1], served=True)
        assert alone["image_hash"][0] == a["image_hash"][i], i
    lat = a["done_ms"] - a["arrival_ms"]
    prev = reqs["final_kind"] == 0
    print(f"served: wall {sa['wall_ms']:.0f} ms, preview p50 {np.percentile(lat[prev], 50):.1f} ms, "
          f"max active {sa['max_active']}; plain: wall {sb['wall_ms']:.0f} ms", flush=True)
    srv.close()
