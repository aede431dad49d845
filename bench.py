"""Benchmark: G rays/s of the MI355X path tracer on the 172,800-triangle
torus workload (1024x1024, 64 spp, depth 6, seed 7), with the roofline of
the closest-hit traversal kernel and the reference CPU renderer timed on this
host beside it.

A step is one full render of the workload through the product path: the
scene compiled by the C++ host (JSON -> Scene::compile -> binned-SAH BVH),
resident on the device, rendered with the reference's progressive pass
schedule (1, 1, 2, 4, 8, 16, 32 samples) into a cleared accumulation buffer.
Rays are counted by the reference's rules (closest-hit rays + shadow rays).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload torus_c2]
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ray-tracer_amd")

WORKLOADS = {
    "torus_c2": "torus_c2.json",
    "cornell_c1": "cornell_c1.json",
    "sphere_field_c3": "sphere_field_c3.json",
}
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic bytes of the FP32 closest-hit kernel: 32 B per node box tested
# (the per-node figure the measurement is defined with), 48 B per triangle
# test, 24 B of ray read (origin + direction) and 16 B of hit written per
# ray (a word for the primitive or plane, t, two barycentrics). The device's
# node records are 32 bytes for both children (half-float boxes), so a box
# tested moves 16 B, and its triangle slots are 36 bytes (v0 and two edges):
# reported beside as the layout's bytes.
NODE_B, TRI_B, RAY_IN_B, HIT_OUT_B, RADIANCE_OUT_B = 32, 48, 24, 16, 12
NODE_LAYOUT_B, TRI_LAYOUT_B = 16, 36


def load_pkg():
    spec = importlib.util.spec_from_file_location("ray_tracer_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ray_tracer_amd"] = mod
    spec.loader.exec_module(mod)
    from ray_tracer_amd import rtg
    return rtg


def render_once(rtg, ctx, dev, acc, cfg, spp, max_pass=32):
    """One progressive render of `spp` samples: the pass schedule of
    renderer.cpp:66-69 in one rtg_render_passes call, as GpuRenderer::render
    issues it when no per-pass callback reads the framebuffer."""
    acc.clear()
    bounds = [0] + [s1 for _, s1 in pass_schedule(spp, max_pass)]
    st, _, _ = ctx.render_passes(dev, cfg, bounds, acc)
    return st.rays + st.shadow_rays


def host_cpus():
    """The host cores this process may use: the affinity set, capped by
    OMP_NUM_THREADS when it is set (the box's CPU share for one GPU), and
    what the machine reports (nproc, model name)."""
    affinity = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else affinity
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": cap,
            "model": model}


def cpu_baseline(scene_text, width, height, spp_full, target_s=12.0, single_s=4.0):
    """The reference renderer (oracle/_ref, built from the reference's own
    sources) on this host's cores, on a bounded sample of the workload:
    whole renders of the full image at a sample count of at most the
    workload's (spp_full), chosen so that the sample takes ~target_s seconds
    on every usable core; then ~single_s seconds on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    cpus = host_cpus()
    if orc.have_ref():
        kind = "reference"

        def run(spp, threads, w=width, h=height):
            r = orc.RefScene(scene_text, width=w, height=h, spp=spp)
            stats = (C.c_uint64 * 5)()
            ms = C.c_double()
            orc.ref().ref_render_full(r.h, threads, 1, 32, None, stats, C.byref(ms))
            return (stats[1] + stats[2]), ms.value / 1e3
    else:
        kind = "port"

        def run(spp, threads, w=width, h=height):
            o = orc.OracleScene(json.loads(scene_text), width=w, height=h)
            t = time.perf_counter()
            _, _, st = o.render(0, spp, threads=threads)
            return st["rays"] + st["shadow_rays"], time.perf_counter() - t

    def sample(threads, budget_s, w, h):
        rays1, s1 = run(1, threads, w, h)
        spp = int(max(1, min(spp_full, round(budget_s / max(s1, 1e-3)))))
        rays, secs, reps = 0, 0.0, 0
        while secs < budget_s * 0.8 and reps < 8:  # whole renders until ~budget_s of CPU work
            r, t = run(spp, threads, w, h) if spp > 1 else (rays1, s1)
            rays, secs, reps = rays + r, secs + t, reps + 1
        return rays, secs, reps, spp

    threads = cpus["threads"]
    rays, secs, reps, spp = sample(threads, target_s, width, height)
    # One thread: a quarter of the image (every fourth row band would change
    # nothing per ray; a smaller image keeps the sample short).
    w1, h1 = max(1, width // 2), max(1, height // 2)
    rays_1, secs_1, reps_1, spp_1 = sample(1, single_s, w1, h1)
    return {"value": rays / secs / 1e9, "unit": "G rays/s", "cores": threads, "kind": kind,
            "sample": f"{reps} x {width}x{height} at {spp} spp (of {spp_full}), progressive passes, "
                      f"{rays} rays in {secs:.2f} s on {threads} threads",
            "host": cpus,
            "single_thread": {"value": rays_1 / secs_1 / 1e9, "unit": "G rays/s", "cores": 1,
                              "sample": f"{reps_1} x {w1}x{h1} at {spp_1} spp, {rays_1} rays in {secs_1:.2f} s"}}


def extend_roofline(prof, traffic, kind="extend"):
    """Roofline of the traversal kernel from a profile: algorithmic bytes per
    launch over its average launch duration (HIP events recorded by each
    launch's dispatch). kind "extend": the path integrator's closest-hit
    kernel (node boxes, primitive tests, ray in + hit out); kind "direct": the
    direct / normals kernel, which traces a sample's closest-hit and shadow
    rays itself with rays and hits in registers (node boxes and primitive
    tests of both, 12 B of radiance out per sample; its recursion frames are
    not counted, so this is a lower bound)."""
    ext = prof[kind]
    n_launch = max(1, ext["launches"])
    if kind == "direct":
        bytes_total = (NODE_B * (prof["closest_visits"] + prof["shadow_visits"]) +
                       TRI_B * (prof["closest_tests"] + prof["shadow_tests"]) + RADIANCE_OUT_B * prof["samples"])
    else:
        bytes_total = (NODE_B * prof["closest_visits"] + TRI_B * prof["closest_tests"] +
                       (RAY_IN_B + HIT_OUT_B) * prof["closest_rays"])
    per_launch = bytes_total / n_launch
    visits = prof["closest_visits"] + (prof["shadow_visits"] if kind == "direct" else 0)
    tests = prof["closest_tests"] + (prof["shadow_tests"] if kind == "direct" else 0)
    layout_per_launch = (bytes_total - (NODE_B - NODE_LAYOUT_B) * visits - (TRI_B - TRI_LAYOUT_B) * tests) / n_launch
    avg_ms = ext["ms"] / n_launch
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    layout_achieved = layout_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # What the HBM itself carries per launch (FETCH_SIZE + WRITE_SIZE counters,
    # scaled by a copy of known size) over the same launch time.
    hbm = traffic / (avg_ms * 1e-3) / 1e9 if traffic and avg_ms > 0 else None
    label = ("k_direct (FP32 direct / normals sample, closest-hit and shadow traversal)" if kind == "direct"
             else "k_extend_dyn (FP32 closest-hit traversal, persistent, lanes refilled)")
    return {"kernel": label, "bound": "hbm", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "hbm_achieved": hbm, "hbm_frac": hbm / HBM_PEAK_GBS if hbm is not None else None,
            "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches": ext["launches"],
            "layout_bytes_per_launch": layout_per_launch, "layout_achieved": layout_achieved,
            "layout_frac": layout_achieved / HBM_PEAK_GBS,
            "node_visits_per_ray": prof["closest_visits"] / max(1, prof["closest_rays"]),
            "tests_per_ray": prof["closest_tests"] / max(1, prof["closest_rays"]),
            "shadow_node_visits_per_ray": prof["shadow_visits"] / max(1, prof["shadow_rays"]),
            "shadow_tests_per_ray": prof["shadow_tests"] / max(1, prof["shadow_rays"])}


def pass_schedule(spp, max_pass=32):
    """Progressive passes (renderer.cpp:66-69): 1, 1, 2, 4, ... capped."""
    out, done = [], 0
    while done < spp:
        k = min(spp - done, max(1, done), max_pass)
        out.append((done, done + k))
        done += k
    return out


def batch_leg(rtg, seed=4, count=64):
    """The batched workload (scenes/make_batch.py): `count` complete renders
    from one seed, each with the reference's progressive pass schedule.
    Plain: one render after another, each pass through rtg_render_pass;
    and one render after another, each in one rtg_render_passes call (what
    GpuRenderer does when no callback reads the passes).
    Batched: pass i of every unfinished render in one rtg_render_batch call.
    The images of all three are compared bitwise."""
    import hashlib
    sys.path.insert(0, os.path.join(PKG, "scenes"))
    import make_batch
    specs = make_batch.build(seed, count)
    t0 = time.perf_counter()
    hs = [rtg.HostScene(json.dumps(j)) for j in specs]
    prep_s = time.perf_counter() - t0
    ctx = rtg.Context(0)
    devs = [ctx.scene(h.desc) for h in hs]
    cfgs = [h.settings(rtg.PRECISION_F32) for h in hs]
    scheds = [pass_schedule(int(h.info()["spp"])) for h in hs]
    accs = [ctx.accum(c.width, c.height) for c in cfgs]
    big = max(range(count), key=lambda i: cfgs[i].width * cfgs[i].height)
    ctx.render_pass(devs[big], cfgs[big], 0, 32, accs[big])  # grow the path buffers before timing
    ctx.render_batch([(devs[big], accs[big], cfgs[big], 0, 32)])
    for a in accs:
        a.clear()
    ctx.synchronize()

    def digests():
        return [hashlib.sha256(b"".join(x.tobytes() for x in a.read())).hexdigest() for a in accs]

    t0 = time.perf_counter()
    rays_p = 0
    for d, a, c, sched in zip(devs, accs, cfgs, scheds):
        for s0, s1 in sched:
            st, _ = ctx.render_pass(d, c, s0, s1, a)
            rays_p += st.rays + st.shadow_rays
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    ref = digests()
    for a in accs:
        a.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    rays_q = 0
    for d, a, c, sched in zip(devs, accs, cfgs, scheds):
        st, _, _ = ctx.render_passes(d, c, [0] + [s1 for _, s1 in sched], a)
        rays_q += st.rays + st.shadow_rays
    ctx.synchronize()
    passes_s = time.perf_counter() - t0
    same_q = digests() == ref
    for a in accs:
        a.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    rays_b, calls = 0, 0
    for p in range(max(len(s) for s in scheds)):
        jobs = [(devs[i], accs[i], cfgs[i]) + scheds[i][p] for i in range(count) if p < len(scheds[i])]
        rays_b += sum(st.rays + st.shadow_rays for st in ctx.render_batch(jobs))
        calls += 1
    ctx.synchronize()
    batched_s = time.perf_counter() - t0
    identical = digests() == ref
    # Every render whole (all its samples) as one job of a single batch call:
    # the 64 renders in one launch sequence; each pixel still adds its
    # samples in order, so the images equal the pass-by-pass ones.
    for a in accs:
        a.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    rays_o = sum(st.rays + st.shadow_rays for st in
                 ctx.render_batch([(devs[i], accs[i], cfgs[i], 0, scheds[i][-1][1]) for i in range(count)]))
    ctx.synchronize()
    one_call_s = time.perf_counter() - t0
    identical = identical and digests() == ref and rays_o == rays_b
    pass_calls = [[(devs[i], accs[i], cfgs[i]) + scheds[i][p] for i in range(count) if p < len(scheds[i])]
                  for p in range(max(len(s) for s in scheds))]
    # The autotuner picks the scheduling knobs (queue / chunk capacity, wide
    # batch passes, lanes) by timing this workload's calls, then the batched
    # workload is timed again under its choice.
    t0 = time.perf_counter()
    tune = ctx.autotune(pass_calls, key=f"batch-seed{seed}-{count}")
    tune["tuning_s"] = round(time.perf_counter() - t0, 2)
    for a in accs:
        a.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    for call in pass_calls:
        ctx.render_batch(call)
    ctx.synchronize()
    tuned_s = time.perf_counter() - t0
    identical = identical and digests() == ref
    for x in accs + devs:
        x.release()
    ctx.close()
    kinds = {}
    for j in specs:
        k = j["name"].split()[0]
        kinds[k] = kinds.get(k, 0) + 1
    return {"workload": f"{count} renders from seed {seed}: " + ", ".join(f"{v} {k}" for k, v in sorted(kinds.items())) +
                        "; 64^2-1024^2, 1-256 spp, progressive passes",
            "paths": int(sum(c.width * c.height * int(h.info()["spp"]) for c, h in zip(cfgs, hs))),
            "rays": int(rays_b), "host_prep_s": round(prep_s, 2),
            "plain_s": round(plain_s, 4), "plain_passes_call_s": round(passes_s, 4),
            "batched_s": round(batched_s, 4), "batch_calls": calls,
            "plain_g_rays_per_s": rays_p / plain_s / 1e9, "plain_passes_call_g_rays_per_s": rays_q / passes_s / 1e9,
            "batched_g_rays_per_s": rays_b / batched_s / 1e9,
            "speedup": plain_s / batched_s, "speedup_vs_passes_call": passes_s / batched_s,
            "batched_one_call_s": round(one_call_s, 4), "batched_one_call_g_rays_per_s": rays_b / one_call_s / 1e9,
            "speedup_one_call": plain_s / one_call_s, "speedup_one_call_vs_passes_call": passes_s / one_call_s,
            "autotune": tune, "batched_tuned_s": round(tuned_s, 4),
            "batched_tuned_g_rays_per_s": rays_b / tuned_s / 1e9,
            "speedup_tuned": plain_s / tuned_s, "speedup_tuned_vs_passes_call": passes_s / tuned_s,
            "identical": identical and same_q and rays_p == rays_b == rays_q}


def serve_leg(rtg, seconds=3.0, loads=(0.5, 0.8, 1.0), seed=5, burst=300, sat_requests=60, tick_paths=0):
    """The served workload over 16 resident scenes of the batched workload:
    80 % previews (128^2, 1-4 spp, depth 4, one pass), 20 % finals (512^2,
    768^2 or 1024^2, 256 spp, the scene's depth, progressive passes capped at
    32). First the saturation rate: `sat_requests` mixed requests arriving at
    once through the plain loop (one request at a time). Then Poisson traces
    of `seconds` at each fraction of that rate in `loads`, each through the
    served loop and the plain loop; then `burst` previews arriving at once.
    Every request's image hash must agree between the loops."""
    import numpy as np
    sys.path.insert(0, os.path.join(PKG, "scenes"))
    import make_batch
    specs = make_batch.build(seed, 16)
    hs = [rtg.HostScene(json.dumps(j)) for j in specs]
    depth = np.array([int(j["render"]["max_depth"]) for j in specs])
    srv = rtg.Server(0, 0, tick_paths)
    for h in hs:
        srv.add_scene(h)
    rng = np.random.default_rng(seed)

    def requests(n, arrivals, final):
        r = np.zeros(n, rtg.REQUEST_DTYPE)
        r["scene"] = rng.integers(0, 16, n)
        r["seed"] = rng.integers(1, 10 ** 9, n)
        r["arrival_ms"] = arrivals
        r["final_kind"] = final
        size = rng.choice([512, 768, 1024], n)
        r["width"] = r["height"] = np.where(final, size, 128)
        r["spp"] = np.where(final, 256, rng.integers(1, 5, n))
        r["max_depth"] = np.where(final, depth[r["scene"]], 4)
        r["max_pass"] = np.where(final, 32, 4)
        return r

    def summary(reqs, r, st):
        lat = r["done_ms"] - r["arrival_ms"]
        prev, fin = reqs["final_kind"] == 0, reqs["final_kind"] == 1
        return {"preview_p50_ms": float(np.percentile(lat[prev], 50)) if prev.any() else None,
                "preview_p99_ms": float(np.percentile(lat[prev], 99)) if prev.any() else None,
                "final_p50_ms": float(np.percentile(lat[fin], 50)) if fin.any() else None,
                "final_first_frame_p50_ms": float(np.percentile((r["first_frame_ms"] - r["arrival_ms"])[fin], 50))
                if fin.any() else None,
                "wall_s": st["wall_ms"] / 1e3, "requests_per_s": len(reqs) / (st["wall_ms"] / 1e3),
                "g_rays_per_s": st["rays"] / st["wall_ms"] / 1e6, "gpu_busy": st["gpu_ms"] / st["wall_ms"],
                "ticks": st["ticks"], "max_active": st["max_active"],
                "peak_footprint_gib": st["peak_footprint_bytes"] / 2 ** 30}

    warm = requests(4, np.zeros(4), np.array([1, 0, 0, 0]))
    warm["width"][0] = warm["height"][0] = 1024
    srv.run(warm, served=True)  # grows the path buffers before timing
    srv.run(warm, served=False)
    out = {"scenes": "16 resident scenes of the batched workload (seed 5)",
           "mix": "80 % previews 128^2 1-4 spp depth 4; 20 % finals 512^2-1024^2 256 spp progressive"}
    # Saturation: the mixed requests all at once; the plain loop's rate is
    # the capacity the load levels below are fractions of.
    sat = requests(sat_requests, np.zeros(sat_requests), rng.random(sat_requests) < 0.2)
    sat_out = {}
    for name, served in (("served", True), ("plain", False)):
        r, st = srv.run(sat, served=served)
        sat_out[name] = summary(sat, r, st)
        sat_out[name]["hash"] = r["image_hash"]
    rate = sat_out["plain"]["requests_per_s"]
    out["saturation"] = {"requests": sat_requests, "plain_requests_per_s": rate,
                         "served_requests_per_s": sat_out["served"]["requests_per_s"],
                         "served_vs_plain": sat_out["served"]["requests_per_s"] / rate,
                         "identical": bool(np.array_equal(sat_out["served"].pop("hash"), sat_out["plain"].pop("hash")))}
    levels = []
    identical = out["saturation"]["identical"]
    for load in loads:
        n = max(5, int(rng.poisson(load * rate * seconds)))
        reqs = requests(n, np.sort(rng.uniform(0.0, seconds * 1e3, n)), rng.random(n) < 0.2)
        lv = {"load": load, "qps": load * rate, "requests": n, "finals": int(reqs["final_kind"].sum())}
        res = {}
        for name, served in (("served", True), ("plain", False)):
            r, st = srv.run(reqs, served=served)
            res[name] = r
            lv[name] = summary(reqs, r, st)
        lv["identical"] = bool(np.array_equal(res["served"]["image_hash"], res["plain"]["image_hash"]))
        identical = identical and lv["identical"]
        levels.append(lv)
    out["levels"] = levels
    pb = requests(burst, np.zeros(burst), np.zeros(burst, bool))
    bursts = {}
    for name, served in (("served", True), ("plain", False)):
        r, st = srv.run(pb, served=served)
        bursts[name] = burst / (st["wall_ms"] / 1e3)
    out["preview_burst"] = {"previews": burst, "served_per_s": bursts["served"], "plain_per_s": bursts["plain"],
                            "ratio": bursts["served"] / bursts["plain"]}
    out["identical"] = identical
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="torus_c2", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--integrator", default=None, choices=["path", "direct", "normals"],
                    help="override the scene's integrator (the headline workload uses path)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-all", action="store_true",
                    help="time every kernel's launches in the timed region (default: the closest-hit kernel's)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 (reference-precision) renders")
    ap.add_argument("--no-serve", action="store_true", help="skip the served workload")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="served workload: trace length per load level")
    ap.add_argument("--serve-loads", default="0.5,0.8,1.0",
                    help="served workload: arrival rates as fractions of the plain loop's saturation rate")
    ap.add_argument("--tick-paths", type=int, default=0,
                    help="served workload: final-render paths per tick (0: tuned per tick)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 times): the reference kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build renders on one GPU; multi-GPU runs are not part of it "
              "(the sample range would shard across GPUs with one FP32 sum+count reduction)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        import torch
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}))
        return
    rtg = load_pkg()
    with open(os.path.join(PKG, "scenes", WORKLOADS[args.workload])) as f:
        text = f.read()
    t0 = time.perf_counter()
    hs = rtg.HostScene(text)
    prep_ms = (time.perf_counter() - t0) * 1e3
    info = hs.info()
    precision = rtg.PRECISION_F64 if args.precision == "f64" else rtg.PRECISION_F32
    cfg = hs.settings(precision)
    if args.integrator:
        cfg.integrator = rtg.INTEGRATORS[args.integrator]
    spp = int(info["spp"])
    ctx = rtg.Context(0)
    dev = ctx.scene(hs.desc)
    acc = ctx.accum(cfg.width, cfg.height)
    for _ in range(args.warmup):
        render_once(rtg, ctx, dev, acc, cfg, spp)
    ctx.synchronize()
    # Only the closest-hit launches are timed in the timed region (the
    # roofline's kernel): each timed launch perturbs the pass a little.
    # FP32 direct runs the weighted ray queue, whose closest-hit kernel
    # (k_dextend) is timed like the path integrator's; FP64 direct and the
    # normals integrator run one kernel per sample.
    queue = cfg.integrator == rtg.INTEGRATORS["direct"] and precision == rtg.PRECISION_F32
    kind = "direct" if cfg.integrator != rtg.INTEGRATORS["path"] and not queue else "extend"
    rtg.set_profiling(ctx, True, kinds=None if args.profile_all else [kind])
    ctx.synchronize()
    t0 = time.perf_counter()
    rays = 0
    for _ in range(args.steps):
        rays += render_once(rtg, ctx, dev, acc, cfg, spp)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    prof = rtg.get_profile(ctx).as_dict()
    prof["samples"] = cfg.width * cfg.height * spp * args.steps
    rtg.set_profiling(ctx, False)
    value = rays / secs / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_extend_{args.workload}.json")
    if kind == "extend" and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = extend_roofline(prof, traffic, kind)
    if queue:
        roofline["kernel"] = "k_dextend (FP32 direct: closest hit of one generation of the weighted ray queue)"
    node_bytes = NODE_B * (prof["closest_visits"] + prof["shadow_visits"])
    node_rate = node_bytes / secs / 1e9
    roofline["concurrent"] = ("direct / normals renders run on one stream; the kernel is alone on the GPU"
                              if kind == "direct" or queue else
                              "the pass schedule runs in one call, its chunks alternating between two streams "
                              "(consecutive passes side by side), each with its any-hit kernels on a third or "
                              "fourth stream, so a launch shares the GPU with up to three other kernels and its "
                              "duration includes that time; roofline_serial is the kernel alone")
    kernel_ms = {k: round(v["ms"] / args.steps, 3) for k, v in prof.items() if isinstance(v, dict) and v["launches"]}
    # The same kernel with every launch of a pass on one stream (not timed
    # for `value`): the traversal kernel's own rate.
    saved = {k: os.environ.get(k) for k in ("RTG_OVERLAP", "RTG_LANES")}
    os.environ["RTG_OVERLAP"] = "0"
    os.environ["RTG_LANES"] = "1"
    rtg.set_profiling(ctx, True)
    for _ in range(2):
        render_once(rtg, ctx, dev, acc, cfg, spp)
    ctx.synchronize()
    sprof = rtg.get_profile(ctx).as_dict()
    sprof["samples"] = cfg.width * cfg.height * spp * 2
    roofline_serial = extend_roofline(sprof, None, kind)  # the traffic file holds the timed render's launches
    roofline_serial["kernel"] = roofline["kernel"]
    roofline_serial["kernel_ms_per_render"] = {k: round(v["ms"] / 2, 3) for k, v in sprof.items()
                                               if isinstance(v, dict) and v["launches"]}
    rtg.set_profiling(ctx, False)
    for k, v in saved.items():
        if v is None:
            del os.environ[k]
        else:
            os.environ[k] = v
    # The reference's own precision: the same renders in FP64 (double
    # arithmetic, the reference's PCG32 stream; bit-identical hits).
    f64 = None
    if precision == rtg.PRECISION_F32 and not args.no_f64:
        cfg64 = hs.settings(rtg.PRECISION_F64)
        cfg64.integrator = cfg.integrator
        render_once(rtg, ctx, dev, acc, cfg64, spp)
        ctx.synchronize()
        t64 = time.perf_counter()
        rays64 = sum(render_once(rtg, ctx, dev, acc, cfg64, spp) for _ in range(2))
        ctx.synchronize()
        s64 = time.perf_counter() - t64
        f64 = {"value": rays64 / s64 / 1e9, "unit": "G rays/s", "ms_per_render": s64 / 2 * 1e3, "renders": 2,
               "dtype": "f64"}
    dev.release()
    acc.release()
    ctx.close()
    batch = None if args.no_batch else batch_leg(rtg)
    served = None if args.no_serve else serve_leg(rtg, seconds=args.serve_seconds,
                                                     loads=tuple(float(x) for x in args.serve_loads.split(",")),
                                                     tick_paths=args.tick_paths)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(text, cfg.width, cfg.height, spp)
    line = {
        "metric": "G rays/s + % HBM roofline; served p50/p99 latency vs plain GPU loop",
        "value": value, "unit": "G rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic (procedural torus scene, seeded RNG)",
        "config": {"workload": f"{args.workload}: {info['triangles']} triangles + {info['spheres']} sphere + "
                               f"{info['planes']} plane, {cfg.width}x{cfg.height}, {spp} spp, depth {cfg.max_depth}, "
                               f"seed {cfg.seed}, progressive passes",
                   "bvh_nodes": info["nodes"], "bvh_depth": info["max_depth"], "scene_prep_ms": round(prep_ms, 1),
                   "rays_per_step": rays // args.steps,
                   "closest_rays_per_step": prof["closest_rays"] // args.steps,
                   "shadow_rays_per_step": (rays - prof["closest_rays"]) // args.steps,
                   "kernel_ms_per_step": kernel_ms},
        "roofline": roofline,
        "roofline_serial": roofline_serial,
        # The whole render's BVH node traffic (32 B per node box tested, by
        # the closest-hit and any-hit kernels together) over its wall time:
        # the node-traffic rate the GPU sustains across concurrent kernels.
        "node_traffic": {"bytes_per_step": node_bytes // args.steps, "achieved": node_rate, "unit": "GB/s",
                         "peak": HBM_PEAK_GBS, "frac": node_rate / HBM_PEAK_GBS,
                         "bytes_per_box": NODE_B, "layout_bytes_per_box": NODE_LAYOUT_B,
                         "layout_frac": node_rate * NODE_LAYOUT_B / NODE_B / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "f64": f64,
        "batched": batch,
        "served": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
