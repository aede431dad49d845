"""CPU tests of the product's host side (librthost.so) and of the C-ABI
libraries: scene reader/validator against the reference's texts, compile +
BVH against the oracle bit for bit, the `rt` CLI's CPU paths, and that both
shared libraries load and export every function their headers declare."""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

import orc
import scenes
from _pkg import rtg

ROOT = scenes.ROOT
META = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_meta.json")))
BAD = json.load(open(os.path.join(ROOT, "tests", "golden", "bad_scenes.json")))
RT = os.path.join(ROOT, "ray-tracer_amd", "bin", "rt")


def header_functions(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:rtg|rth|rts|rt)_[a-z0-9_]+)\s*\(", text)) - {"rt_frame_fn", "rt_done_fn"})


@pytest.mark.parametrize("lib,header", [("librtg.so", "rtg.h"), ("librthost.so", "rt_host.h")])
def test_library_exports_every_declared_symbol(lib, header):
    L = C.CDLL(os.path.join(ROOT, "ray-tracer_amd", "lib", lib))
    names = header_functions(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    if header == "rtg.h":
        assert sorted(rtg.rtg_symbols()) == names  # the binding covers the header
    else:
        assert sorted(rtg.host_symbols()) == names


def test_rtg_reports_no_device_cleanly():
    lib = rtg.load_rtg()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert lib.rtg_open(0, C.byref(h)) == -6
    assert b"no HIP device" in lib.rtg_last_error(None)


ALL_SCENES = [("asset", n) for n in sorted(os.listdir(scenes.ASSET_DIR))] + [("cfg", "cornell_c1.json"),
                                                                             ("cfg", "torus_c2.json")]


@pytest.mark.parametrize("kind,name", ALL_SCENES)
def test_host_compile_matches_oracle(kind, name):
    j = scenes.load_asset(name) if kind == "asset" else scenes.load(name)
    h = rtg.HostScene(json.dumps(j))
    o = orc.OracleScene(j)
    a, b = orc.arrays(h.desc), orc.arrays(o.desc)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert bytes(h.desc.camera) == bytes(o.desc.camera)
    assert bytes(h.desc.background) == bytes(o.desc.background)
    mats = lambda d: bytes((rtg.Material * d.n_materials).from_address(C.addressof(d.materials.contents)))  # noqa
    assert mats(h.desc) == mats(o.desc)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_host_bvh_matches_oracle_on_random_scenes(seed):
    j = scenes.random_scene_json(300, 600, seed, seed == 2)
    a, b = orc.arrays(rtg.HostScene(json.dumps(j)).desc), orc.arrays(orc.OracleScene(j).desc)
    assert all(np.array_equal(a[k], b[k]) for k in a)


def test_dump_scene_matches_reference():
    for fn, want in META["dump_scene_sha256"].items():
        text = open(os.path.join(scenes.ASSET_DIR, fn)).read()
        got = rtg.dump_scene(text)
        assert hashlib.sha256(got.encode()).hexdigest() == want, fn
        # and the round trip is lossless (tests/test_scene_io.cpp:37-53)
        assert rtg.dump_scene(got) == got


def test_dump_of_escapes_and_extreme_numbers_matches_reference():
    """Unicode escapes (surrogate pairs), control characters, -0, 1e-300 and
    the largest double survive the reader and print as the reference prints
    them."""
    cases = json.load(open(os.path.join(ROOT, "tests", "golden", "dump_scenes.json")))
    assert len(cases) == len(META["dump_scene_texts"])
    for text, want in zip(cases, META["dump_scene_texts"]):
        got = rtg.dump_scene(text)
        assert got == want
        assert rtg.dump_scene(got) == got


def test_validation_messages_match_reference():
    for text, want in zip(BAD, META["bad_scene_messages"]):
        assert rtg.validate(text) == want, text[:80]


def test_defaults_and_overrides():
    h = rtg.HostScene("{}")
    s = h.settings()
    assert (s.width, s.height, s.max_depth, s.rr_min_depth, s.seed, s.light_sampling) == (640, 360, 8, 3, 1, 1)
    assert h.desc.n_prims == 0 and h.desc.n_nodes == 0
    h2 = rtg.HostScene(json.dumps(scenes.load("cornell_c1.json")), width=32, height=16, spp=3, seed=9,
                       integrator="direct", use_bvh=False, light_sampling=False)
    s2 = h2.settings(rtg.PRECISION_F64)
    assert (s2.width, s2.height, s2.seed, s2.integrator, s2.use_bvh, s2.light_sampling, s2.precision) == \
        (32, 16, 9, 1, 0, 0, 1)
    assert h2.info()["spp"] == 3


def test_rt_cli_cpu_paths():
    scene = os.path.join(scenes.SCENE_DIR, "torus_c2.json")
    r = subprocess.run([RT, scene, "--validate"], capture_output=True, timeout=60)
    assert r.returncode == 0 and b"scene OK: 3 objects" in r.stderr
    r = subprocess.run([RT, scene, "--dump-scene", "--spp", "5"], capture_output=True, timeout=60)
    d = json.loads(r.stdout)
    assert r.returncode == 0 and d["render"]["spp"] == 5 and d["objects"][0]["segments_v"] == 180
    r = subprocess.run([RT, "-", "--stream"], input=b"not json", capture_output=True, timeout=60)
    assert r.returncode == 2 and r.stdout.startswith(b"ERROR ") and b"JSON parse error" in r.stdout
    r = subprocess.run([RT, scene, "--integrator", "magic"], capture_output=True, timeout=60)
    assert r.returncode == 2 and b"--integrator" in r.stderr
    r = subprocess.run([RT, scene, "--bogus"], capture_output=True, timeout=60)
    assert r.returncode == 1 and b"unknown option" in r.stderr
    r = subprocess.run([RT], capture_output=True, timeout=60)
    assert r.returncode == 1


NODE_SCENES = [("cornell_c1.json", None), ("torus_c2.json", None), ("sphere_field_c3.json", None), (None, 1),
               (None, 2), (None, 3), ("far", 1), ("huge", 3)]


def node_scene_json(name, seed):
    """A shipped scene, a random scene (tests/test_bvh.cpp:11-30), or a
    random scene moved 1e4 units from the origin ("far") or scaled up past
    what half floats hold even in the node frame ("huge": its boxes become
    infinite there, which still bounds them)."""
    if name == "far":
        return scenes.translated(scenes.random_scene_json(300, 600, seed, False), (1.0e4, -3.0e3, 2.5e3))
    if name == "huge":
        j = scenes.random_scene_json(300, 600, seed, False)
        f = 1.0e12
        for o in j["objects"]:
            for k in ("center", "v0", "v1", "v2", "point"):
                if k in o:
                    o[k] = [c * f for c in o[k]]
            if "radius" in o:
                o["radius"] *= f
        return j
    return scenes.load(name) if name else scenes.random_scene_json(300, 600, seed, seed == 2)


@pytest.mark.parametrize("name,seed", NODE_SCENES)
def test_fp32_node_boxes_contain_the_double_boxes(name, seed):
    """Every child box of the device's FP32 node records (half floats in the
    scene's node frame) contains the double box of the reference BVH node it
    stands for (include/raytracer/geometry/bvh.hpp:20-26) and is the tightest
    half box that does, child refs point at the right records or leaf ranges,
    and the reported split axis is the node's."""
    j = node_scene_json(name, seed)
    hs = rtg.HostScene(json.dumps(j))
    d = hs.desc
    nodes = np.frombuffer((rtg.BvhNode * d.n_nodes).from_address(C.addressof(d.nodes.contents)),
                          dtype=np.dtype([("lo", "<f8", 3), ("hi", "<f8", 3), ("offset", "<u4"), ("count", "<u2"),
                                          ("axis", "u1"), ("pad", "u1")]))
    rec = rtg.node_records(d)
    origin, e = rtg.node_frame(d)
    interior = np.flatnonzero(nodes["count"] == 0)
    assert len(rec) == len(interior) > 0
    record_of = np.full(len(nodes), -1, np.int64)
    record_of[interior] = np.arange(len(interior))
    left, right = interior + 1, nodes["offset"][interior].astype(np.int64)
    half = rec[:, :12].astype(np.float16)
    assert np.array_equal(half.astype(np.float32), rec[:, :12])  # the planes are half floats
    world = lambda h: origin.astype(np.float64) + h.astype(np.float64) * 2.0 ** e  # exact in double
    with np.errstate(invalid="ignore", over="ignore"):
        for child, lo, hi in ((left, half[:, 0:3], half[:, 3:6]), (right, half[:, 6:9], half[:, 9:12])):
            assert (world(lo) <= nodes["lo"][child]).all() and (world(hi) >= nodes["hi"][child]).all()
            # outward rounding is tight: the next half inward no longer contains the box
            assert (world(np.nextafter(lo, np.float16(np.inf))) > nodes["lo"][child]).all()
            assert (world(np.nextafter(hi, np.float16(-np.inf))) < nodes["hi"][child]).all()
    if name == "far":  # the frame follows the scene: nothing is infinite, boxes stay tight
        assert np.isfinite(half).all() and abs(origin[0] - 1.0e4) < 100
    if name == "huge":
        assert e == 20 and np.isinf(half).any()
    refs = rec[:, 12:15].view(np.uint32)
    assert np.array_equal(refs[:, 2], nodes["axis"][interior])
    for col, child in ((0, left), (1, right)):
        r = refs[:, col]
        leaf = nodes["count"][child] > 0
        assert np.array_equal((r >> 31) == 1, leaf)
        assert np.array_equal(r[~leaf], record_of[child[~leaf]])
        assert np.array_equal(r[leaf] & 0x1FFFFFF, nodes["offset"][child[leaf]])
        assert np.array_equal(((r[leaf] >> 25) & 63) + 1, nodes["count"][child[leaf]])
    hs.close()
