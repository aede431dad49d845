"""The oracle pinned: against the reference's own outputs (golden vectors
written by tests/golden/make_golden.py from the reference build), against the
known answers of the reference's unit tests (tests/test_*.cpp), and against
closed forms. CPU only."""
import ctypes as C
import hashlib
import json
import math
import os

import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
META = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_meta.json")))


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------- golden ---
def test_rng_streams_match_reference():
    for (s, p, k), want in zip(GOLD["rng_grid"], GOLD["rng_streams"]):
        buf = (C.c_uint32 * 8)()
        orc.orc().orc_rng_stream(int(s), int(p), int(k), buf, 8)
        assert list(buf) == list(want)


@pytest.mark.parametrize("key,loader,name", [
    ("asset_cornell_box", scenes.load_asset, "cornell_box.json"),
    ("asset_first_light", scenes.load_asset, "first_light.json"),
    ("asset_materials", scenes.load_asset, "materials.json"),
    ("asset_sphere_field", scenes.load_asset, "sphere_field.json"),
    ("asset_soft_shadows", scenes.load_asset, "soft_shadows.json"),
    ("asset_mesh_bvh", scenes.load_asset, "mesh_bvh.json"),
    ("cfg_cornell_c1", scenes.load, "cornell_c1.json"),
    ("cfg_torus_c2", scenes.load, "torus_c2.json"),
])
def test_compile_and_bvh_match_reference(key, loader, name):
    o = orc.OracleScene(loader(name))
    a = orc.arrays(o.desc)
    for k, v in a.items():
        assert sha(v) == META[key][k], f"{key}.{k}"
        if f"{key}_{k}" in GOLD:
            assert np.array_equal(v, GOLD[f"{key}_{k}"])
    info = META[key]["info"]
    assert o.desc.n_nodes == info["nodes"] and o.c.bvh.max_depth == info["max_depth"]
    assert o.c.bvh.leaves == info["leaves"] and o.c.degenerate_skipped == info["degenerate_skipped"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_scene_bvh_matches_reference(seed):
    o = orc.OracleScene(scenes.random_scene_json(300, 600, seed, seed == 2))
    a = orc.arrays(o.desc)
    assert np.array_equal(a["nodes"], GOLD[f"random_{seed}_nodes"])
    assert np.array_equal(a["prim_indices"], GOLD[f"random_{seed}_prim_indices"])
    assert sha(a["prims"]) == META[f"random_{seed}"]["prims"]


@pytest.mark.parametrize("seed", [1, 2])
def test_hits_match_reference(seed):
    o = orc.OracleScene(scenes.random_scene_json(300, 600, seed, seed == 2))
    rays = GOLD[f"hits_random_{seed}_rays"]
    h, nv, pt = o.trace(rays, 0, 1)
    assert np.array_equal(h.view(np.uint8).reshape(len(h), -1), GOLD[f"hits_random_{seed}_closest"])
    assert (nv, pt) == (META[f"hits_random_{seed}"]["node_visits"], META[f"hits_random_{seed}"]["prim_tests"])
    hb, _, _ = o.trace(rays, 0, 0)
    assert np.array_equal(hb.view(np.uint8).reshape(len(h), -1), GOLD[f"hits_random_{seed}_brute"])
    r2 = rays.copy()
    r2[:, 7] = 5.0
    ha, _, _ = o.trace(r2, 1, 1)
    assert np.array_equal(ha.view(np.uint8).reshape(len(h), -1), GOLD[f"hits_random_{seed}_any5"])


def test_torus_hits_match_reference():
    o = orc.OracleScene(scenes.load("torus_c2.json"))
    rays = GOLD["hits_torus_rays"]
    h, nv, pt = o.trace(rays, 0, 1)
    assert np.array_equal(h.view(np.uint8).reshape(len(h), -1), GOLD["hits_torus_closest"])
    assert (nv, pt) == (META["hits_torus"]["node_visits"], META["hits_torus"]["prim_tests"])


FB_CASES = [("fb_cornell_c1", "cornell_c1.json", None, False), ("fb_torus_c2", "torus_c2.json", None, False),
            ("fb_first_light", "first_light.json", None, True), ("fb_materials_direct", "materials.json", "direct", True),
            ("fb_cornell_glass_metal", "cornell_glass_metal.json", None, True),
            ("fb_sphere_field", "sphere_field.json", None, True), ("fb_mirrors_normals", "mirrors.json", "normals", True)]


@pytest.mark.parametrize("key,name,integ,asset", FB_CASES)
def test_framebuffer_sums_match_reference(key, name, integ, asset):
    m = META[key]
    j = scenes.load_asset(name) if asset else scenes.load(name)
    o = orc.OracleScene(j, width=m["w"], height=m["h"], integrator=integ)
    s, c, st = o.render(0, m["spp"], threads=4)
    assert np.array_equal(c, GOLD[key + "_count"])
    assert np.array_equal(s.view(np.uint64), GOLD[key + "_sum"].view(np.uint64)), "sums differ in the last bit"
    assert [st["camera_samples"], st["rays"], st["shadow_rays"], st["rr_terminated"], st["invalid_samples"]] == m["stats"]


DEEP = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_direct_deep.npz"))


@pytest.mark.parametrize("name,w,h,spp,depth", [("mirror_corridor.json", 12, 8, 2, 1000),
                                                ("mirror_corridor.json", 12, 8, 2, 200),
                                                ("cornell_glass_metal.json", 10, 10, 2, 40)])
def test_deep_direct_matches_reference(name, w, h, spp, depth):
    """direct_trace far past 31 bounces (max_depth up to 1000)."""
    if name == "mirror_corridor.json":
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)) as f:
            j = json.load(f)
    else:
        j = scenes.load_asset(name)
    key = f"{name[:-5]}_{depth}"
    o = orc.OracleScene(j, width=w, height=h, integrator="direct", max_depth=depth)
    s, c, st = o.render(0, spp, threads=4)
    assert np.array_equal(c, DEEP[key + "_count"])
    assert np.array_equal(s.view(np.uint64), DEEP[key + "_sum"].view(np.uint64))
    assert [st["camera_samples"], st["rays"], st["shadow_rays"], st["rr_terminated"], st["invalid_samples"]] == \
        [int(x) for x in DEEP[key + "_stats"]]


def test_camera_rays_match_reference():
    cam = GOLD["camera_desc"]
    camera = rtg.Camera()
    arr = lambda v: (C.c_double * 3)(*v)  # noqa: E731
    orc.orc().orc_camera_setup(arr(cam[0:3]), arr(cam[3:6]), arr(cam[6:9]), cam[9], cam[10], cam[11], int(cam[12]),
                               int(cam[13]), C.byref(camera))
    L = C.CDLL(orc.ORC_PATH)
    L.orc_camera_generate.argtypes = [C.POINTER(rtg.Camera), C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_rng_for_sample.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
    for k, (px, want) in enumerate(zip(GOLD["camera_px"], GOLD["camera_rays"])):
        rng = (C.c_uint64 * 2)()
        L.orc_rng_for_sample(rng, 5, 0, k)
        o, d = (C.c_double * 3)(), (C.c_double * 3)()
        L.orc_camera_generate(C.byref(camera), int(px[0]), int(px[1]), px[2], px[3], rng, o, d)
        assert list(o) + list(d) == list(want)


def test_tone_map_and_srgb_match_reference():
    L = orc.orc()
    for v, want in zip(GOLD["srgb_in"], GOLD["srgb_out"]):
        assert L.orc_linear_to_srgb8(float(v)) == want
    for op in range(3):
        for i, v in enumerate(GOLD["srgb_in"]):
            a, b = (C.c_double * 3)(v, v * 0.5, v * 2), (C.c_double * 3)()
            L.orc_tone_map(a, op, 1.3, b)
            assert list(b) == list(GOLD["tonemap_out"][op, i])


# ------------------------------------------------------- known answers ---
def one_object_scene(obj, mat=None):
    return {"materials": {"m": mat or {"type": "lambertian"}}, "objects": [dict(obj, material="m")]}


def hit(o, orig, d, tmin=1e-9, tmax=math.inf):
    rec = orc.OrcHit()
    ok = orc.orc().orc_intersect(C.byref(o.desc), 1, (C.c_double * 3)(*orig), (C.c_double * 3)(*d), tmin, tmax,
                                 C.byref(rec), None, None)
    return ok, rec


def test_sphere_known_answers():  # tests/test_geometry.cpp:12-46
    o = orc.OracleScene(one_object_scene({"type": "sphere", "center": [0, 0, -5], "radius": 1}))
    ok, r = hit(o, (0, 0, 0), (0, 0, -1))
    assert ok and r.t == pytest.approx(4) and r.front_face and r.normal[2] == pytest.approx(1)
    assert not hit(o, (0, 2, 0), (0, 0, -1))[0]
    assert not hit(o, (0, 0, 0), (0, 0, 1))[0]
    ok, r = hit(o, (0, 0, -5), (0, 0, -1))
    assert ok and r.t == pytest.approx(1) and not r.front_face and r.normal[2] == pytest.approx(1)
    assert not hit(o, (0, 0, 0), (0, 0, -1), tmax=3.9)[0]
    ok, r = hit(o, (0, 0, 0), (0, 0, -2))
    assert ok and r.t == pytest.approx(2)
    tiny = orc.OracleScene(one_object_scene({"type": "sphere", "center": [0, 0, -1e4], "radius": 1e-3}))
    ok, r = hit(tiny, (0, 0, 0), (0, 0, -1))
    assert ok and abs(r.t - (1e4 - 1e-3)) <= 1e-12 * (1e4 - 1e-3)


def test_triangle_and_plane_known_answers():  # tests/test_geometry.cpp:60-117
    tri = orc.OracleScene(one_object_scene({"type": "triangle", "v0": [0, 0, 0], "v1": [1, 0, 0], "v2": [0, 1, 0]}))
    ok, r = hit(tri, (0.25, 0.25, 1), (0, 0, -1))
    assert ok and r.t == pytest.approx(1) and list(r.normal) == [0, 0, 1] and r.front_face
    assert not hit(tri, (0.6, 0.6, 1), (0, 0, -1))[0] and not hit(tri, (-0.1, 0.5, 1), (0, 0, -1))[0]
    ok, r = hit(tri, (0.2, 0.2, -1), (0, 0, 1))
    assert ok and not r.front_face and list(r.normal) == [0, 0, -1]
    pl = orc.OracleScene(one_object_scene({"type": "plane", "point": [0, -1, 0], "normal": [0, 1, 0]}))
    s = 1 / math.sqrt(2)
    ok, r = hit(pl, (0, 1, 0), (s, -s, 0))
    assert ok and r.p[1] == pytest.approx(-1) and r.p[0] == pytest.approx(2) and r.is_plane
    assert not hit(pl, (0, 1, 0), (1, 0, 0))[0]
    ok, r = hit(pl, (0, -3, 0), (0, 1, 0))
    assert ok and not r.front_face and list(r.normal) == [0, -1, 0]


def test_slab_known_answers():  # tests/test_math.cpp:96-117 (incl. the 0*inf NaN and -0.0 cases)
    L = orc.orc()
    L.orc_box_hit.restype = C.c_int
    L.orc_box_hit.argtypes = [C.POINTER(C.c_double)] * 4 + [C.c_double, C.c_double]
    v = lambda *x: (C.c_double * 3)(*x)  # noqa: E731
    lo, hi = v(-1, -1, -1), v(1, 1, 1)
    cases = [((0, 0, 5), (0, 0, -1), 0, math.inf, True), ((0, 0, 5), (0, 0, 1), 0, math.inf, False),
             ((2, 0, 5), (0, 0, -1), 0, math.inf, False), ((0, 0, 5), (0, 0, -1), 0, 3.9, False),
             ((0, 0, 0), (0.3, 0.2, 0.1), 0, math.inf, True), ((1, 0, 5), (0, 0, -1), 0, math.inf, True),
             ((1, 0, 5), (-0.0, 0, -1), 0, math.inf, True), ((-1, 0, 5), (0.0, 0, -1), 0, math.inf, True),
             ((-1, 0, 5), (-0.0, 0, -1), 0, math.inf, True), ((1.001, 0, 5), (-0.0, 0, -1), 0, math.inf, False),
             ((-5, -5, -5), (1, 1, 1), 0, math.inf, True)]
    for o, d, t0, t1, want in cases:
        assert bool(L.orc_box_hit(lo, hi, v(*o), v(*d), t0, t1)) == want, (o, d)


def test_srgb_known_answers():  # tests/test_render.cpp:268-278
    L = orc.orc()
    assert [L.orc_linear_to_srgb8(v) for v in (0, 1, 0.5, -1, 10)] == [0, 255, 188, 0, 255]


def test_camera_centre_and_corner_rays():  # tests/test_math.cpp:170-191
    o = orc.OracleScene({"camera": {"look_from": [0, 0, 0], "look_at": [0, 0, -1], "vfov": 90},
                         "render": {"width": 200, "height": 100}})
    c = o.desc.camera
    assert (c.w.x, c.w.y, c.w.z) == pytest.approx((0, 0, 1)) and (c.u.x, c.u.y, c.u.z) == pytest.approx((1, 0, 0))
    L = C.CDLL(orc.ORC_PATH)
    L.orc_camera_generate.argtypes = [C.POINTER(rtg.Camera), C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = (C.c_uint64 * 2)(1, 1)
    oo, d = (C.c_double * 3)(), (C.c_double * 3)()
    L.orc_camera_generate(C.byref(c), 100, 50, 0, 0, rng, oo, d)
    assert list(d) == pytest.approx([0, 0, -1], abs=1e-12)
    L.orc_camera_generate(C.byref(c), 0, 0, 0, 0, rng, oo, d)
    want = np.array([-2, 1, -1]) / math.sqrt(6)
    assert list(d) == pytest.approx(list(want), abs=1e-12)


def bvh(boxes, max_leaf=4):
    b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
    n = len(b)
    nodes = (rtg.BvhNode * max(2 * n, 1))()
    pi = (C.c_uint32 * max(n, 1))()
    nn, npi, st = C.c_uint32(), C.c_uint32(), orc.OrcBvhStats()
    assert orc.orc().orc_bvh_build(b.ctypes.data_as(C.POINTER(C.c_double)), n, max_leaf, nodes, C.byref(nn), pi,
                                   C.byref(npi), C.byref(st)) == 0
    ok = orc.orc().orc_bvh_validate(nodes, nn.value, pi, b.ctypes.data_as(C.POINTER(C.c_double)), n)
    return nodes[:nn.value], list(pi)[:npi.value], st, ok


def test_bvh_edge_cases():  # tests/test_bvh.cpp:79-109
    nodes, pi, st, ok = bvh(np.zeros((0, 6)))
    assert len(nodes) == 0 and pi == []
    nodes, pi, st, ok = bvh([[-1, -1, -1, 1, 1, 1]])
    assert len(nodes) == 1 and nodes[0].count == 1 and ok
    nodes, pi, st, ok = bvh([[0, 0, 0, 1, 1, 1]] * 1000)
    assert ok and st.max_leaf_size <= 4 and sorted(pi) == list(range(1000))
    inf = math.inf
    nodes, pi, st, ok = bvh([[0, 0, 0, 1, 1, 1], [inf, inf, inf, -inf, -inf, -inf], [2, 0, 0, 3, 1, 1]])
    assert ok and sorted(pi) == [0, 2]


def test_degenerate_triangles_dropped():  # tests/test_bvh.cpp:110-121
    o = orc.OracleScene({"materials": {"m": {"type": "lambertian"}}, "objects": [
        {"type": "triangle", "v0": [0, 0, 0], "v1": [1, 0, 0], "v2": [0, 1, 0], "material": "m"},
        {"type": "mesh", "source": "icosphere", "scale": [1, 1e-30, 1], "material": "m"}]})
    assert o.c.degenerate_skipped > 0


def test_point_light_inverse_square():  # tests/test_render.cpp:153-171
    j = {"materials": {"white": {"type": "lambertian", "albedo": [1, 1, 1]}},
         "objects": [{"type": "plane", "point": [0, 0, 0], "normal": [0, 1, 0], "material": "white"}],
         "lights": [{"position": [0, 2, 0], "intensity": [8, 8, 8]}]}
    o = orc.OracleScene(j)
    out = (C.c_double * 3)()
    assert orc.orc().orc_sample_direct_light_at(C.byref(o.desc), (C.c_double * 3)(0, 1, 0),
                                                (C.c_double * 3)(0, -1, 0), 1, out)
    assert out[0] == pytest.approx(2 / math.pi)
    j["objects"].append({"type": "sphere", "center": [0, 1, 0], "radius": 0.2, "material": "white"})
    o2 = orc.OracleScene(j)
    assert orc.orc().orc_sample_direct_light_at(C.byref(o2.desc), (C.c_double * 3)(0, 0.5, 0),
                                                (C.c_double * 3)(0, -1, 0), 1, out)
    assert list(out) == [0, 0, 0]


def test_sphere_light_over_plane_closed_form():  # tests/test_render.cpp:89-99 (smaller run)
    j = scenes.sphere_light_over_plane(True)
    o = orc.OracleScene(j)
    s, c, st = o.render(0, 1024, threads=4)
    assert float((s[..., 1] / c).mean()) == pytest.approx(0.2, rel=0.02)
    assert st["invalid_samples"] == 0


def test_progressive_schedule_and_determinism():  # tests/test_render.cpp:62-87, 211-221
    o = orc.OracleScene(scenes.load("cornell_c1.json"), width=12, height=12)
    a, ca, _ = o.render(0, 8, threads=1)
    b = np.zeros_like(a)
    cb = np.zeros_like(ca)
    for s0, s1 in [(0, 1), (1, 2), (2, 4), (4, 8)]:
        s, c, _ = o.render(s0, s1, threads=3)
        b += s  # different pass boundaries, same per-pixel sample order
        cb += c
    assert np.array_equal(ca, cb)
    np.testing.assert_allclose(a, b, rtol=1e-12)
