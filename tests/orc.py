"""Test-side bindings of the oracle (oracle/liborc.so) and, where it was
built, of the reference renderer itself (oracle/_ref/librtref.so).

Test infrastructure only. `scene_from_json` turns a scene file into the
oracle's object list with the defaults of the reference's scene reader
(src/io/scene_io.cpp:101-310) and compiles it with the oracle.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import threading

import numpy as np

from _pkg import rtg  # noqa: E402  (the product's ctypes structures)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "librtref.so")

P = C.POINTER
VP = C.c_void_p


class OrcObject(C.Structure):
    _fields_ = [("kind", C.c_int32), ("material", C.c_uint32), ("a", C.c_double * 3), ("b", C.c_double * 3),
                ("c", C.c_double * 3), ("scalar", C.c_double), ("mesh_source", C.c_int32),
                ("subdivisions", C.c_int32), ("segments_u", C.c_int32), ("segments_v", C.c_int32),
                ("smooth", C.c_int32), ("pad", C.c_int32), ("major_radius", C.c_double),
                ("minor_radius", C.c_double), ("rotate_y", C.c_double), ("scale", C.c_double * 3),
                ("translate", C.c_double * 3)]


class OrcBvhStats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("leaves", C.c_uint64), ("max_depth", C.c_uint64),
                ("max_leaf_size", C.c_uint64), ("sah_cost", C.c_double)]


class OrcCompiled(C.Structure):
    _fields_ = [("prims", P(rtg.Primitive)), ("n_prims", C.c_uint32), ("is_sampled_light", P(C.c_uint8)),
                ("planes", P(rtg.Plane)), ("n_planes", C.c_uint32), ("lights", P(rtg.LightRef)),
                ("n_lights", C.c_uint32), ("nodes", P(rtg.BvhNode)), ("n_nodes", C.c_uint32),
                ("prim_indices", P(C.c_uint32)), ("n_prim_indices", C.c_uint32),
                ("degenerate_skipped", C.c_uint64), ("bvh", OrcBvhStats)]


class OrcHit(C.Structure):
    _fields_ = [("t", C.c_double), ("p", C.c_double * 3), ("normal", C.c_double * 3), ("gn", C.c_double * 3),
                ("front_face", C.c_int32), ("material", C.c_uint32), ("primitive", C.c_uint32),
                ("is_plane", C.c_int32)]


class RefOverrides(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("rr_min_depth", C.c_int32), ("integrator", C.c_int32), ("light_sampling", C.c_int32),
                ("use_bvh", C.c_int32), ("seed", C.c_int64), ("force_segments_v_path", C.c_int32),
                ("pad", C.c_int32)]


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise RuntimeError("oracle/liborc.so missing: run make -C oracle")
        L = C.CDLL(ORC_PATH)
        sig = {
            "orc_rng_stream": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint32), C.c_int]),
            "orc_bvh_build": (C.c_int, [P(C.c_double), C.c_uint32, C.c_int, P(rtg.BvhNode), P(C.c_uint32),
                                        P(C.c_uint32), P(C.c_uint32), P(OrcBvhStats)]),
            "orc_bvh_validate": (C.c_int, [P(rtg.BvhNode), C.c_uint32, P(C.c_uint32), P(C.c_double), C.c_uint32]),
            "orc_prim_bounds": (None, [P(rtg.Primitive), P(C.c_double)]),
            "orc_intersect": (C.c_int, [P(rtg.SceneDesc), C.c_int, P(C.c_double), P(C.c_double), C.c_double,
                                        C.c_double, P(OrcHit), P(C.c_uint64), P(C.c_uint64)]),
            "orc_trace_batch": (None, [P(rtg.SceneDesc), C.c_int, C.c_int, VP, C.c_uint64, VP, P(C.c_uint64),
                                       P(C.c_uint64)]),
            "orc_camera_setup": (None, [P(C.c_double), P(C.c_double), P(C.c_double), C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int, P(rtg.Camera)]),
            "orc_render_rows": (None, [P(rtg.SceneDesc), P(rtg.RenderSettings), C.c_uint32, C.c_uint32, C.c_int,
                                       C.c_int, VP, VP, P(rtg.Stats)]),
            "orc_sample_radiance_rows": (None, [P(rtg.SceneDesc), P(rtg.RenderSettings), C.c_uint32, C.c_uint32,
                                                C.c_int, C.c_int, VP, P(rtg.Stats)]),
            "orc_sample_direct_light_at": (C.c_int, [P(rtg.SceneDesc), P(C.c_double), P(C.c_double), C.c_uint64,
                                                     P(C.c_double)]),
            "orc_tone_map": (None, [P(C.c_double), C.c_int, C.c_double, P(C.c_double)]),
            "orc_linear_to_srgb8": (C.c_uint8, [C.c_double]),
            "orc_resolve": (None, [VP, VP, C.c_int, C.c_int, C.c_int, C.c_double, VP]),
            "orc_compile": (C.c_int, [P(OrcObject), C.c_uint32, P(rtg.Material), C.c_uint32, C.c_uint32,
                                      P(OrcCompiled)]),
            "orc_compiled_free": (None, [P(OrcCompiled)]),
            "orc_make_torus": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, P(P(C.c_double)),
                                         P(P(C.c_double)), P(C.c_uint32), P(P(C.c_uint32)), P(C.c_uint32)]),
            "orc_make_icosphere": (C.c_int, [C.c_int, P(P(C.c_double)), P(P(C.c_double)), P(C.c_uint32),
                                             P(P(C.c_uint32)), P(C.c_uint32)]),
            "orc_free": (None, [VP]),
            "orc_random_scene_objects": (C.c_uint32, [C.c_int, C.c_int, C.c_uint64, C.c_int, P(OrcObject)]),
            "orc_ray_set": (None, [C.c_uint64, C.c_uint64, P(C.c_double), P(C.c_double), VP]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _orc = L
    return _orc


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        sig = {
            "ref_last_error": (C.c_char_p, []),
            "ref_scene_from_json": (VP, [C.c_char_p, P(RefOverrides)]),
            "ref_scene_free": (None, [VP]),
            "ref_scene_desc": (P(rtg.SceneDesc), [VP]),
            "ref_scene_info": (None, [VP, P(C.c_uint64)]),
            "ref_scene_camera_desc": (None, [VP, P(C.c_double)]),
            "ref_camera_rays": (None, [VP, VP, C.c_uint64, C.c_uint64, VP]),
            "ref_trace_batch": (None, [VP, C.c_int, C.c_int, VP, C.c_uint64, VP, P(C.c_uint64), P(C.c_uint64)]),
            "ref_render_rows": (None, [VP, C.c_uint32, C.c_uint32, C.c_int, C.c_int, VP, VP, P(C.c_uint64)]),
            "ref_render_moments": (None, [VP, C.c_uint32, C.c_uint32, C.c_int, C.c_int, VP, VP, VP,
                                          P(C.c_uint64)]),
            "ref_render_full": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, VP, P(C.c_uint64), P(C.c_double)]),
            "ref_rng_stream": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint32), C.c_int]),
            "ref_bvh_build": (C.c_int, [P(C.c_double), C.c_uint32, C.c_int, P(rtg.BvhNode), P(C.c_uint32),
                                        P(C.c_uint32), P(C.c_uint32)]),
            "ref_make_torus": (C.c_uint32, [C.c_double, C.c_double, C.c_int, C.c_int, VP, VP, VP]),
            "ref_linear_to_srgb8": (C.c_uint8, [C.c_double]),
            "ref_tone_map": (None, [P(C.c_double), C.c_int, C.c_double, P(C.c_double)]),
            "ref_sample_direct_light": (C.c_int, [VP, P(C.c_double), P(C.c_double), C.c_uint64, P(C.c_double)]),
            "ref_dump_scene": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint64]),
            "ref_validate": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint64]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _ref = L
    return _ref


# ---------------------------------------------------------------------------
# JSON -> oracle scene, with the reference reader's defaults.

def _v3(d, key, default):
    v = d.get(key, default)
    if isinstance(v, (int, float)):
        return [float(v)] * 3
    return [float(x) for x in v]


MAT_KIND = {"lambertian": 0, "diffuse": 0, "metal": 1, "dielectric": 2, "glass": 2, "emissive": 3, "light": 3}


def _material(m):
    k = MAT_KIND[m["type"]]
    r = rtg.Material()
    r.kind = k
    if k == 0:
        a = _v3(m, "albedo", 0.5)
        r.c0 = rtg.Vec3(*a)
        r.c1 = rtg.Vec3(*_v3(m, "albedo2", a))
        r.p0 = float(m.get("checker_scale", 0))
    elif k == 1:
        r.c0 = rtg.Vec3(*_v3(m, "albedo", 0.9))
        r.p0 = float(m.get("roughness", 0))
    elif k == 2:
        r.c0 = rtg.Vec3(*_v3(m, "tint", 1.0))
        r.p0 = float(m.get("ior", 1.5))
    else:
        r.c0 = rtg.Vec3(*_v3(m, "color", 1.0))
        r.p0 = float(m.get("intensity", 1))
        r.two_sided = 1 if m.get("two_sided", False) else 0
    return r


def _object(o, mat_index):
    r = OrcObject()
    r.material = mat_index[o["material"]]
    t = o["type"]
    if t == "sphere":
        r.kind = 0
        r.a[:] = _v3(o, "center", 0)
        r.scalar = float(o.get("radius", 1))
    elif t == "plane":
        r.kind = 1
        r.a[:] = _v3(o, "point", 0)
        r.b[:] = _v3(o, "normal", [0, 1, 0])
    elif t == "triangle":
        r.kind = 2
        r.a[:], r.b[:], r.c[:] = _v3(o, "v0", 0), _v3(o, "v1", 0), _v3(o, "v2", 0)
    elif t == "quad":
        r.kind = 3
        r.a[:], r.b[:], r.c[:] = _v3(o, "corner", 0), _v3(o, "u", 0), _v3(o, "v", 0)
    elif t == "box":
        r.kind = 4
        r.a[:], r.b[:] = _v3(o, "min", 0), _v3(o, "max", 0)
        r.scalar = float(o.get("rotate_y", 0))
    elif t == "mesh":
        r.kind = 5
        src = o.get("source", "icosphere")
        assert src in ("icosphere", "torus"), "the oracle has no OBJ reader"
        r.mesh_source = 1 if src == "torus" else 0
        r.subdivisions = int(o.get("subdivisions", 3))
        r.major_radius = float(o.get("major_radius", 1))
        r.minor_radius = float(o.get("minor_radius", 0.35))
        r.segments_u = int(o.get("segments", 48))
        r.segments_v = int(o.get("segments_v", 0))
        r.smooth = 1 if o.get("smooth", True) else 0
        r.scale[:] = _v3(o, "scale", [1, 1, 1])
        r.rotate_y = float(o.get("rotate_y", 0))
        r.translate[:] = _v3(o, "translate", [0, 0, 0])
    else:
        raise ValueError(t)
    return r


def settings_from_json(j, **ov):
    R = dict(width=640, height=360, spp=64, max_depth=8, rr_min_depth=3, integrator="path", light_sampling=True,
             seed=1, tone_map="aces", exposure=1.0, use_bvh=True)
    R.update(j.get("render", {}))
    R.update({k: v for k, v in ov.items() if v is not None})
    return R


class OracleScene:
    """A scene compiled by the oracle, plus its SceneDesc view."""

    def __init__(self, j, **overrides):
        if isinstance(j, str):
            j = json.loads(j)
        self.json = j
        self.R = settings_from_json(j, **overrides)
        L = orc()
        mats = list(j.get("materials", {}).items())
        mat_index = {n: i for i, (n, _) in enumerate(mats)}
        self.materials = (rtg.Material * max(len(mats), 1))(*[_material(m) for _, m in mats])
        objs = [_object(o, mat_index) for o in j.get("objects", [])]
        self._objs = (OrcObject * max(len(objs), 1))(*objs)
        lights = j.get("lights", [])
        self.point_lights = (rtg.PointLight * max(len(lights), 1))(
            *[rtg.PointLight(rtg.Vec3(*_v3(l, "position", 0)), rtg.Vec3(*_v3(l, "intensity", 10.0))) for l in lights])
        self.c = OrcCompiled()
        rc = L.orc_compile(self._objs, len(objs), self.materials, len(mats), len(lights), C.byref(self.c))
        assert rc == 0, rc
        cam = j.get("camera", {})
        camera = rtg.Camera()
        lf = (C.c_double * 3)(*_v3(cam, "look_from", [0, 0, 0]))
        la = (C.c_double * 3)(*_v3(cam, "look_at", [0, 0, -1]))
        up = (C.c_double * 3)(*_v3(cam, "vup", [0, 1, 0]))
        L.orc_camera_setup(lf, la, up, float(cam.get("vfov", 40)), float(cam.get("aperture", 0)),
                           float(cam.get("focus_dist", 0)), int(self.R["width"]), int(self.R["height"]),
                           C.byref(camera))
        bg = j.get("background", {})
        b = rtg.Background()
        kind = bg.get("type", "gradient")
        b.kind = 1 if kind == "solid" else 0
        b.top = rtg.Vec3(*_v3(bg, "color" if kind == "solid" else "top", [0.5, 0.7, 1.0]))
        b.bottom = rtg.Vec3(*_v3(bg, "bottom", [1, 1, 1]))
        b.intensity = float(bg.get("intensity", 1))
        d = rtg.SceneDesc()
        d.prims, d.n_prims = self.c.prims, self.c.n_prims
        d.is_sampled_light = self.c.is_sampled_light
        d.planes, d.n_planes = self.c.planes, self.c.n_planes
        d.materials, d.n_materials = C.cast(self.materials, P(rtg.Material)), len(mats)
        d.point_lights, d.n_point_lights = C.cast(self.point_lights, P(rtg.PointLight)), len(lights)
        d.lights, d.n_lights = self.c.lights, self.c.n_lights
        d.nodes, d.n_nodes = self.c.nodes, self.c.n_nodes
        d.prim_indices, d.n_prim_indices = self.c.prim_indices, self.c.n_prim_indices
        d.camera, d.background = camera, b
        self.desc = d
        d._owner = self  # the arrays live in the oracle's compiled scene

    def __del__(self):
        try:
            orc().orc_compiled_free(C.byref(self.c))
        except Exception:
            pass

    def settings(self, precision=1):
        R = self.R
        return rtg.RenderSettings(int(R["width"]), int(R["height"]), int(R["max_depth"]), int(R["rr_min_depth"]),
                                  rtg.INTEGRATORS[R["integrator"]], 1 if R["light_sampling"] else 0,
                                  1 if R["use_bvh"] else 0, precision, int(R["seed"]))

    def render(self, s0, s1, threads=8):
        """Oracle framebuffer sum/count for samples [s0, s1), rows split over threads."""
        cfg = self.settings()
        w, h = cfg.width, cfg.height
        s = np.zeros((h, w, 3))
        c = np.zeros((h, w), np.uint32)
        stats = [rtg.Stats() for _ in range(threads)]
        rows = np.linspace(0, h, threads + 1).astype(int)

        def work(k):
            orc().orc_render_rows(C.byref(self.desc), C.byref(cfg), s0, s1, int(rows[k]), int(rows[k + 1]),
                                  C.c_void_p(s.ctypes.data), C.c_void_p(c.ctypes.data), C.byref(stats[k]))

        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        [t.start() for t in ts]
        [t.join() for t in ts]
        tot = {k: sum(getattr(st, k) for st in stats) for k, _ in rtg.Stats._fields_ if k != "elapsed_ms"}
        return s, c, tot

    def trace(self, rays, mode=0, use_bvh=1):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
        hits = np.zeros(len(rays), rtg.HIT_DTYPE)
        nv, pt = C.c_uint64(0), C.c_uint64(0)
        orc().orc_trace_batch(C.byref(self.desc), mode, use_bvh, C.c_void_p(rays.ctypes.data), len(rays),
                              C.c_void_p(hits.ctypes.data), C.byref(nv), C.byref(pt))
        return hits, nv.value, pt.value


class RefScene:
    """The reference renderer's own compiled scene (oracle/_ref)."""

    def __init__(self, text, force_segments_v_path=False, **ov):
        L = ref()
        o = RefOverrides(-1, -1, -1, -1, -1, -1, -1, -1, -1, 1 if force_segments_v_path else 0, 0)
        for k, v in ov.items():
            if v is None:
                continue
            if k == "integrator":
                v = rtg.INTEGRATORS[v]
            setattr(o, k, int(v))
        if not isinstance(text, str):
            text = json.dumps(text)
        self.h = L.ref_scene_from_json(text.encode(), C.byref(o))
        if not self.h:
            raise RuntimeError(L.ref_last_error().decode())
        self.desc = L.ref_scene_desc(self.h).contents
        self.desc._owner = self  # the arrays live in the C++ scene

    def __del__(self):
        try:
            ref().ref_scene_free(self.h)
        except Exception:
            pass

    def info(self):
        a = (C.c_uint64 * 8)()
        ref().ref_scene_info(self.h, a)
        return dict(zip(["triangles", "spheres", "planes", "degenerate_skipped", "lights", "nodes", "leaves",
                         "max_depth"], list(a)))

    def trace(self, rays, mode=0, use_bvh=1):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
        hits = np.zeros(len(rays), rtg.HIT_DTYPE)
        nv, pt = C.c_uint64(0), C.c_uint64(0)
        ref().ref_trace_batch(self.h, mode, use_bvh, C.c_void_p(rays.ctypes.data), len(rays),
                              C.c_void_p(hits.ctypes.data), C.byref(nv), C.byref(pt))
        return hits, nv.value, pt.value

    def render_rows(self, s0, s1, w, h, threads=8):
        s = np.zeros((h, w, 3))
        c = np.zeros((h, w), np.uint32)
        rows = np.linspace(0, h, threads + 1).astype(int)
        stats = [(C.c_uint64 * 5)() for _ in range(threads)]

        def work(k):
            ref().ref_render_rows(self.h, s0, s1, int(rows[k]), int(rows[k + 1]), C.c_void_p(s.ctypes.data),
                                  C.c_void_p(c.ctypes.data), stats[k])

        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        [t.start() for t in ts]
        [t.join() for t in ts]
        return s, c, [sum(st[i] for st in stats) for i in range(5)]

    def render_moments(self, s0, s1, w, h, threads=8):
        """Per-pixel sum, sum of squares and count of the reference's samples
        [s0, s1) (ref_render_moments); rows are handed out one at a time."""
        s = np.zeros((h, w, 3))
        q = np.zeros((h, w, 3))
        c = np.zeros((h, w), np.uint32)
        stats = [(C.c_uint64 * 5)() for _ in range(threads)]
        nxt = [0]
        lock = threading.Lock()

        def work(k):
            while True:
                with lock:
                    y = nxt[0]
                    nxt[0] += 1
                if y >= h:
                    return
                ref().ref_render_moments(self.h, s0, s1, y, y + 1, C.c_void_p(s.ctypes.data),
                                         C.c_void_p(q.ctypes.data), C.c_void_p(c.ctypes.data), stats[k])

        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        [t.start() for t in ts]
        [t.join() for t in ts]
        return s, q, c, [sum(st[i] for st in stats) for i in range(5)]


def ray_set(desc, n, seed=7):
    """The seeded ray set over twice the BVH root box (orc_ray_set)."""
    root = desc.nodes[0]
    lo = (C.c_double * 3)(root.lo.x, root.lo.y, root.lo.z)
    hi = (C.c_double * 3)(root.hi.x, root.hi.y, root.hi.z)
    rays = np.zeros((n, 8))
    orc().orc_ray_set(seed, n, lo, hi, C.c_void_p(rays.ctypes.data))
    return rays


def ids_digest(hits):
    """SHA-256 of the hit primitive ids and of the t bits of a hit array."""
    import hashlib
    return (hashlib.sha256(np.ascontiguousarray(hits["prim"]).tobytes()).hexdigest(),
            hashlib.sha256(np.ascontiguousarray(hits["t"]).view(np.uint64).tobytes()).hexdigest())


# ---------------------------------------------------------------------------
def arrays(desc):
    """numpy views of a SceneDesc's arrays (for bit-level comparisons)."""
    def view(ptr, n, ctype):
        if n == 0:
            return np.zeros(0, np.uint8)
        return np.frombuffer((ctype * n).from_address(C.addressof(ptr.contents)), dtype=np.uint8).copy()

    return {
        "prims": view(desc.prims, desc.n_prims, rtg.Primitive),
        "planes": view(desc.planes, desc.n_planes, rtg.Plane),
        "lights": view(desc.lights, desc.n_lights, rtg.LightRef),
        "nodes": view(desc.nodes, desc.n_nodes, rtg.BvhNode),
        "prim_indices": view(desc.prim_indices, desc.n_prim_indices, C.c_uint32),
        "sampled": view(desc.is_sampled_light, desc.n_prims, C.c_uint8),
    }
