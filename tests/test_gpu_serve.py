"""The served render loop (rts_*): requests over resident scenes replayed by
the tick scheduler and by the plain loop. A request's image must not depend
on what it was batched with: both loops must give every request the same
final sums and counts (hashed), the same ray counts and the full sample
count, and a request rendered alone must hash the same."""
import json
import os
import sys

import numpy as np
import pytest

import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(scenes.SCENE_DIR))
import make_batch  # noqa: E402


def fnv1a(*arrays):
    """FNV-1a over 64-bit words, each element widened (the server's image hash)."""
    h = 1469598103934665603
    for a in arrays:
        a = np.ascontiguousarray(a).reshape(-1)
        words = a.view(np.uint64) if a.dtype.itemsize == 8 else a.astype(np.uint64)
        for w in words.tolist():
            h = ((h ^ w) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture(scope="module")
def resident():
    js = make_batch.build(seed=11, count=4) + [scenes.load("cornell_c1.json")]
    hs = [rtg.HostScene(json.dumps(j)) for j in js]
    return hs


def trace(n_prev=14, n_final=3, seed=3):
    rng = np.random.default_rng(seed)
    r = np.zeros(n_prev + n_final, rtg.REQUEST_DTYPE)
    r["scene"] = rng.integers(0, 5, len(r))
    r["seed"] = rng.integers(1, 10 ** 6, len(r))
    if n_final:  # finals among the previews (request 0 stays a preview)
        r["final_kind"][rng.choice(np.arange(1, len(r)), n_final, replace=False)] = 1
    r["width"] = np.where(r["final_kind"] == 1, 80, 40)
    r["height"] = np.where(r["final_kind"] == 1, 48, 32)
    r["spp"] = np.where(r["final_kind"] == 1, 37, rng.integers(1, 5, len(r)))
    r["max_depth"] = np.where(r["final_kind"] == 1, 6, 4)
    r["max_pass"] = np.where(r["final_kind"] == 1, 8, 4)
    r["arrival_ms"] = np.sort(rng.uniform(0, 60, len(r)))
    return r


def check_complete(reqs, res):
    assert (res["spp_done"] == reqs["spp"]).all()
    assert (res["done_ms"] >= res["arrival_ms"]).all()
    assert (res["first_frame_ms"] <= res["done_ms"]).all()
    # progressive frames: passes of 1, 1, 2, 4, ... capped at max_pass
    for q, o in zip(reqs, res):
        done, frames = 0, 0
        while done < q["spp"]:
            done += min(q["spp"] - done, max(1, done), q["max_pass"])
            frames += 1
        assert o["frames"] == frames


def test_served_equals_plain(resident):
    srv = rtg.Server(0, 0, 20000)  # small ticks: finals are sliced into many pixel ranges
    for h in resident:
        srv.add_scene(h)
    reqs = trace()
    a, sa = srv.run(reqs, served=True)
    b, sb = srv.run(reqs, served=False)
    check_complete(reqs, a)
    check_complete(reqs, b)
    assert np.array_equal(a["image_hash"], b["image_hash"])
    assert np.array_equal(a["rays"], b["rays"]) and np.array_equal(a["shadow_rays"], b["shadow_rays"])
    assert sa["ticks"] > 0 and sa["max_active"] >= 1 and sb["max_active"] == 1
    print("served", sa, "plain", sb, flush=True)
    srv.close()


def test_request_alone_hashes_the_same(resident, ctx):
    """One preview rendered alone through rtg_render_pass vs inside a mix."""
    srv = rtg.Server(0, 0, 0)
    for h in resident:
        srv.add_scene(h)
    reqs = trace(seed=9)
    mixed, _ = srv.run(reqs, served=True)
    srv.close()
    q = reqs[0]
    j = make_batch.build(seed=11, count=4)[q["scene"]] if q["scene"] < 4 else scenes.load("cornell_c1.json")
    j["render"].update(width=int(q["width"]), height=int(q["height"]), seed=int(q["seed"]),
                       max_depth=int(q["max_depth"]))
    hs = rtg.HostScene(json.dumps(j))
    dev = ctx.scene(hs.desc)
    acc = ctx.accum(int(q["width"]), int(q["height"]))
    ctx.render_pass(dev, hs.settings(rtg.PRECISION_F32), 0, int(q["spp"]), acc)
    s, c = acc.read()
    acc.release()
    dev.release()
    assert fnv1a(s, c) == int(mixed["image_hash"][0])


def test_admission_by_memory_budget(resident, ctx):
    """Admission counts the HBM footprint: resident scenes, the context's
    path state and queues, and the admitted requests' accumulation buffers.
    A budget that leaves room for one final's buffer: requests queue, none is
    lost, the footprint never exceeds the budget while more than one request
    is live, and the images are those of an unconstrained run."""
    reqs = trace(seed=5)
    big = rtg.Server(0, 0, 0)
    small = rtg.Server(0, 0, 0)
    for h in resident:
        big.add_scene(h)
        small.add_scene(h)
    a, sa = big.run(reqs, served=True)
    _, warm = small.run(reqs, served=True)  # grows the path state to what this trace needs
    devs = [ctx.scene(h.desc) for h in resident]
    scene_bytes = sum(d.nbytes() for d in devs)
    for d in devs:
        d.release()
    assert warm["fixed_bytes"] > scene_bytes > 0  # the scenes, and the path state and queues
    one = int(reqs["width"].max() * reqs["height"].max() * 28)
    budget = warm["fixed_bytes"] + one
    small.set_budget(budget)
    b, sb = small.run(reqs, served=True)
    check_complete(reqs, b)
    assert np.array_equal(a["image_hash"], b["image_hash"])
    assert sb["peak_accum_bytes"] <= one and sb["peak_footprint_bytes"] <= budget
    assert sa["peak_accum_bytes"] > one  # unconstrained, more than one final's worth was live at once
    print("fixed", warm["fixed_bytes"], "peak footprint", sb["peak_footprint_bytes"], flush=True)
    big.close()
    small.close()


def test_bad_requests_are_rejected(resident):
    srv = rtg.Server(0, 0, 0)
    srv.add_scene(resident[0])
    r = trace(2, 0)
    r["scene"] = 0
    bad = r.copy()
    bad["scene"][1] = 7
    with pytest.raises(rtg.RtgError, match="no such scene"):
        srv.run(bad)
    bad = r.copy()
    bad["max_depth"][0] = 65
    with pytest.raises(rtg.RtgError, match="max_depth"):
        srv.run(bad)
    bad = r[::-1].copy()
    bad["arrival_ms"] = [5.0, 1.0]
    with pytest.raises(rtg.RtgError, match="sorted"):
        srv.run(bad)
    srv.close()
