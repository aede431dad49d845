"""Writes the golden fixtures of tests/golden/ from the REFERENCE renderer
itself (oracle/_ref/librtref.so, compiled from the reference's sources by
oracle/Makefile). Run in the container that has the reference checkout:

    make -C oracle && python tests/golden/make_golden.py

`--messages` rewrites only the scene-format fixtures (validation messages of
bad_scenes.json, dumps of dump_scenes.json) in reference_meta.json;
`--direct-deep` only reference_direct_deep.npz (the direct integrator far
past 31 bounces).
Each fixture is data only: inputs and the reference's outputs.
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import orc  # noqa: E402
import scenes  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def ray_set(seed, n, extent, axis_every=50):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for i in range(0, n, axis_every):
        d[i] = (0.0, 0.0, 1.0 if i % 100 else -1.0)
    r = np.zeros((n, 8))
    r[:, :3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, 1e-6, np.inf
    return r


def scene_format(L, meta):
    """Validation messages and normalised dumps of the scene reader."""
    bad = json.load(open(os.path.join(HERE, "bad_scenes.json")))
    msgs = []
    for text in bad:
        buf = C.create_string_buffer(4096)
        L.ref_validate(text.encode(), buf, len(buf))
        msgs.append(buf.value.decode())
    meta["bad_scene_messages"] = msgs
    dumps = []
    for text in json.load(open(os.path.join(HERE, "dump_scenes.json"))):
        buf = C.create_string_buffer(1 << 20)
        L.ref_dump_scene(text.encode(), buf, len(buf))
        dumps.append(buf.value.decode())
    meta["dump_scene_texts"] = dumps


# The direct integrator recursing deep: a corridor of two facing mirrors
# (tests/golden/mirror_corridor.json) at max_depth 1000 and 200, and glass
# and metal (two-way splits) at max_depth 40.
DIRECT_DEEP = [("mirror_corridor.json", 12, 8, 2, 1000), ("mirror_corridor.json", 12, 8, 2, 200),
               ("cornell_glass_metal.json", 10, 10, 2, 40)]


def deep_scene(name):
    if name == "mirror_corridor.json":
        with open(os.path.join(HERE, name)) as f:
            return json.load(f)
    return scenes.load_asset(name)


def direct_deep():
    assert orc.have_ref(), "oracle/_ref/librtref.so missing: run make -C oracle in the reference container"
    out = {}
    for name, w, h_, spp, depth in DIRECT_DEEP:
        r = orc.RefScene(json.dumps(deep_scene(name)), width=w, height=h_, integrator="direct", max_depth=depth)
        s, c, st = r.render_rows(0, spp, w, h_)
        key = f"{name[:-5]}_{depth}"
        out[key + "_sum"], out[key + "_count"] = s, c
        out[key + "_stats"] = np.array([int(x) for x in st], np.uint64)
    path = os.path.join(HERE, "reference_direct_deep.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


def messages_only():
    assert orc.have_ref(), "oracle/_ref/librtref.so missing: run make -C oracle in the reference container"
    path = os.path.join(HERE, "reference_meta.json")
    meta = json.load(open(path))
    scene_format(orc.ref(), meta)
    with open(path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("updated", path)


def main():
    assert orc.have_ref(), "oracle/_ref/librtref.so missing: run make -C oracle in the reference container"
    L = orc.ref()
    out, meta = {}, {}

    # 1. RNG streams: first 8 outputs of Rng::for_sample on a grid.
    grid = np.array([(s, p, k) for s in (0, 1, 7, 99, 2**40 + 3) for p in (0, 1, 12345, 2**32 - 1)
                     for k in (0, 1, 63, 100000)], dtype=np.uint64)
    streams = np.zeros((len(grid), 8), np.uint32)
    for i, (s, p, k) in enumerate(grid):
        buf = (C.c_uint32 * 8)()
        L.ref_rng_stream(int(s), int(p), int(k), buf, 8)
        streams[i] = list(buf)
    out["rng_grid"], out["rng_streams"] = grid, streams

    # 2. Compiled scenes + BVHs: exact bytes for small scenes, hashes for big ones.
    names = ["cornell_box.json", "first_light.json", "materials.json", "sphere_field.json", "soft_shadows.json",
             "mesh_bvh.json"]
    for name in names:
        r = orc.RefScene(json.dumps(scenes.load_asset(name)))
        a = orc.arrays(r.desc)
        key = "asset_" + name[:-5]
        meta[key] = {k: sha(v) for k, v in a.items()}
        meta[key]["info"] = r.info()
        if r.desc.n_prims < 2000:
            for k, v in a.items():
                out[f"{key}_{k}"] = v
    for name in ["cornell_c1.json", "torus_c2.json"]:
        r = orc.RefScene(json.dumps(scenes.load(name)))
        a = orc.arrays(r.desc)
        key = "cfg_" + name[:-5]
        meta[key] = {k: sha(v) for k, v in a.items()}
        meta[key]["info"] = r.info()
        if r.desc.n_prims < 2000:
            for k, v in a.items():
                out[f"{key}_{k}"] = v
    for seed in (1, 2, 3):
        r = orc.RefScene(json.dumps(scenes.random_scene_json(300, 600, seed, seed == 2)))
        a = orc.arrays(r.desc)
        meta[f"random_{seed}"] = {k: sha(v) for k, v in a.items()}
        for k in ("nodes", "prim_indices"):
            out[f"random_{seed}_{k}"] = a[k]
    # The segments_v compile path agrees with Scene::compile where both apply.
    soft = json.dumps(scenes.load_asset("soft_shadows.json"))
    a1 = orc.arrays(orc.RefScene(soft).desc)
    a2 = orc.arrays(orc.RefScene(soft, force_segments_v_path=True).desc)
    assert all(np.array_equal(a1[k], a2[k]) for k in a1), "segments_v path differs from Scene::compile"

    # 3. Closest / any hits of seeded rays (Scene::intersect / occluded).
    for seed in (1, 2):
        r = orc.RefScene(json.dumps(scenes.random_scene_json(300, 600, seed, seed == 2)))
        rays = ray_set(seed * 100, 4096, 15.0)
        h, nv, pt = r.trace(rays, 0, 1)
        hb, _, _ = r.trace(rays, 0, 0)
        r2 = rays.copy()
        r2[:, 7] = 5.0
        ho, _, _ = r.trace(r2, 1, 1)
        out[f"hits_random_{seed}_rays"] = rays
        out[f"hits_random_{seed}_closest"] = h.view(np.uint8).reshape(len(h), -1)
        out[f"hits_random_{seed}_brute"] = hb.view(np.uint8).reshape(len(h), -1)
        out[f"hits_random_{seed}_any5"] = ho.view(np.uint8).reshape(len(h), -1)
        meta[f"hits_random_{seed}"] = {"node_visits": nv, "prim_tests": pt}
    torus = orc.RefScene(json.dumps(scenes.load("torus_c2.json")))
    rays = ray_set(7, 8192, 2.0, axis_every=97)
    h, nv, pt = torus.trace(rays, 0, 1)
    out["hits_torus_rays"] = rays
    out["hits_torus_closest"] = h.view(np.uint8).reshape(len(h), -1)
    meta["hits_torus"] = {"node_visits": nv, "prim_tests": pt}

    # 3b. The 2^24-ray torus set (orc_ray_set, seed 7) through Scene::intersect:
    # digests of every hit id and t, hit count and traversal counters.
    torus = orc.RefScene(json.dumps(scenes.torus_only()))
    rays = orc.ray_set(torus.desc, 1 << 24, seed=7)
    hits = np.zeros(len(rays), orc.rtg.HIT_DTYPE)
    parts = np.array_split(np.arange(len(rays)), 8)
    counters = [(C.c_uint64(0), C.c_uint64(0)) for _ in parts]

    def run(k):
        idx = parts[k]
        sub = np.ascontiguousarray(rays[idx])
        out = np.zeros(len(sub), orc.rtg.HIT_DTYPE)
        L.ref_trace_batch(torus.h, 0, 1, sub.ctypes.data_as(C.c_void_p), len(sub), out.ctypes.data_as(C.c_void_p),
                          C.byref(counters[k][0]), C.byref(counters[k][1]))
        hits[idx] = out

    import threading
    ts = [threading.Thread(target=run, args=(k,)) for k in range(len(parts))]
    [t.start() for t in ts]
    [t.join() for t in ts]
    id_sha, t_sha = orc.ids_digest(hits)
    meta["ray_set_2p24"] = {"ids_sha256": id_sha, "t_sha256": t_sha, "hits": int((hits["flags"] & 1).sum()),
                            "node_visits": int(sum(c[0].value for c in counters)),
                            "prim_tests": int(sum(c[1].value for c in counters)),
                            "first_rays_sha256": hashlib.sha256(rays[:4096].tobytes()).hexdigest()}
    out["ray_set_head_hits"] = hits[:65536].view(np.uint8).reshape(65536, -1)

    # 4. Framebuffer sums (render_tile's sample loop through radiance()).
    for name, w, h_, spp, ov in [("cornell_c1.json", 16, 16, 4, {}), ("torus_c2.json", 12, 12, 2, {})]:
        r = orc.RefScene(json.dumps(scenes.load(name)), width=w, height=h_, **ov)
        s, c, st = r.render_rows(0, spp, w, h_)
        key = "fb_" + name[:-5]
        out[key + "_sum"], out[key + "_count"] = s, c
        meta[key] = {"stats": [int(x) for x in st], "w": w, "h": h_, "spp": spp}
    for name, w, h_, spp, integ in [("first_light.json", 16, 12, 3, None), ("materials.json", 16, 8, 2, "direct"),
                                    ("cornell_glass_metal.json", 12, 12, 2, None),
                                    ("sphere_field.json", 16, 10, 2, None), ("mirrors.json", 12, 8, 2, "normals")]:
        r = orc.RefScene(json.dumps(scenes.load_asset(name)), width=w, height=h_, integrator=integ)
        s, c, st = r.render_rows(0, spp, w, h_)
        key = "fb_" + name[:-5] + ("_" + integ if integ else "")
        out[key + "_sum"], out[key + "_count"] = s, c
        meta[key] = {"stats": [int(x) for x in st], "w": w, "h": h_, "spp": spp, "integrator": integ,
                     "scene": name}

    # 5. Camera rays (Camera::generate) of the thin-lens sphere field camera.
    r = orc.RefScene(json.dumps(scenes.load_asset("sphere_field.json")))
    px = np.array([[i, j, sx, sy] for i, j in ((0, 0), (399, 225), (799, 449)) for sx, sy in ((0, 0), (0.5, 0.25))],
                  dtype=np.float64)
    rays_out = np.zeros((len(px), 6))
    L.ref_camera_rays(r.h, px.ctypes.data_as(C.c_void_p), len(px), 5, rays_out.ctypes.data_as(C.c_void_p))
    cam = (C.c_double * 14)()
    L.ref_scene_camera_desc(r.h, cam)
    out["camera_px"], out["camera_rays"], out["camera_desc"] = px, rays_out, np.array(list(cam))

    # 6. Tone map and sRGB encode.
    vals = np.concatenate([np.linspace(-0.5, 1.5, 257), [0.0031308, 0.5, 1e9]])
    out["srgb_in"] = vals
    out["srgb_out"] = np.array([L.ref_linear_to_srgb8(float(v)) for v in vals], np.uint8)
    tm = np.zeros((3, len(vals), 3))
    for op in range(3):
        for i, v in enumerate(vals):
            a = (C.c_double * 3)(v, v * 0.5, v * 2)
            b = (C.c_double * 3)()
            L.ref_tone_map(a, op, 1.3, b)
            tm[op, i] = list(b)
    out["tonemap_out"] = tm

    # 7. Scene JSON round trip (scene_to_json) of every shipped scene.
    dumps = {}
    for fn in sorted(os.listdir(scenes.ASSET_DIR)):
        text = open(os.path.join(scenes.ASSET_DIR, fn)).read()
        buf = C.create_string_buffer(1 << 22)
        L.ref_dump_scene(text.encode(), buf, len(buf))
        dumps[fn] = hashlib.sha256(buf.value).hexdigest()
    meta["dump_scene_sha256"] = dumps

    # 8. Validation messages of the reference for malformed scenes, and
    # normalised dumps of scenes with escapes and extreme numbers.
    scene_format(L, meta)

    np.savez_compressed(OUT, **out)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", OUT, os.path.getsize(OUT), "bytes;", len(out), "arrays")


if __name__ == "__main__":
    if "--messages" in sys.argv[1:]:
        messages_only()
    elif "--direct-deep" in sys.argv[1:]:
        direct_deep()
    else:
        main()
