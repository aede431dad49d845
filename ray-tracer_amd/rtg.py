"""ctypes binding of the device back end's C ABI (include/rtg.h, librtg.so)
and of the host library (include/rt_host.h, librthost.so).

The structures mirror rtg.h field for field. Loading fails loudly when a
library is missing: there is no CPU fallback for any entry point here.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")


class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class BvhNode(C.Structure):
    _fields_ = [("lo", Vec3), ("hi", Vec3), ("offset", C.c_uint32), ("count", C.c_uint16),
                ("axis", C.c_uint8), ("pad", C.c_uint8)]


class Primitive(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("material", C.c_uint32), ("has_vertex_normals", C.c_uint32),
                ("pad", C.c_uint32), ("v0", Vec3), ("v1", Vec3), ("v2", Vec3),
                ("n0", Vec3), ("n1", Vec3), ("n2", Vec3)]


class Plane(C.Structure):
    _fields_ = [("point", Vec3), ("normal", Vec3), ("material", C.c_uint32), ("pad", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("two_sided", C.c_uint32), ("c0", Vec3), ("c1", Vec3), ("p0", C.c_double)]


class PointLight(C.Structure):
    _fields_ = [("position", Vec3), ("intensity", Vec3)]


class LightRef(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("index", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("origin", Vec3), ("corner", Vec3), ("du", Vec3), ("dv", Vec3), ("u", Vec3), ("v", Vec3),
                ("w", Vec3), ("lens_radius", C.c_double)]


class Background(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("top", Vec3), ("bottom", Vec3),
                ("intensity", C.c_double)]


class SceneDesc(C.Structure):
    _fields_ = [("prims", C.POINTER(Primitive)), ("n_prims", C.c_uint32),
                ("is_sampled_light", C.POINTER(C.c_uint8)),
                ("planes", C.POINTER(Plane)), ("n_planes", C.c_uint32),
                ("materials", C.POINTER(Material)), ("n_materials", C.c_uint32),
                ("point_lights", C.POINTER(PointLight)), ("n_point_lights", C.c_uint32),
                ("lights", C.POINTER(LightRef)), ("n_lights", C.c_uint32),
                ("nodes", C.POINTER(BvhNode)), ("n_nodes", C.c_uint32),
                ("prim_indices", C.POINTER(C.c_uint32)), ("n_prim_indices", C.c_uint32),
                ("camera", Camera), ("background", Background)]


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("max_depth", C.c_int32),
                ("rr_min_depth", C.c_int32), ("integrator", C.c_int32), ("light_sampling", C.c_int32),
                ("use_bvh", C.c_int32), ("precision", C.c_int32), ("seed", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("camera_samples", C.c_uint64), ("rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("rr_terminated", C.c_uint64), ("invalid_samples", C.c_uint64), ("node_visits", C.c_uint64),
                ("prim_tests", C.c_uint64), ("elapsed_ms", C.c_double), ("redone_chunks", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Job(C.Structure):
    """rtg_job: one pass (samples [s0, s1) of pixels [p0, p1)) of one render."""
    _fields_ = [("scene", C.c_void_p), ("acc", C.c_void_p), ("settings", RenderSettings),
                ("s0", C.c_uint32), ("s1", C.c_uint32), ("p0", C.c_uint32), ("p1", C.c_uint32)]


class Tuning(C.Structure):
    """rtg_tuning: scheduling knobs (0 = built-in default)."""
    _fields_ = [("chunk_paths", C.c_uint32), ("wide_paths", C.c_uint32), ("lanes", C.c_uint32), ("pad", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class TuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("default_ms", C.c_double), ("best_ms", C.c_double), ("trials", C.c_uint32),
                ("pad", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_double), ("prim", C.c_uint32), ("flags", C.c_uint32)]


HIT_DTYPE = np.dtype([("t", "<f8"), ("prim", "<u4"), ("flags", "<u4")])


class TraceOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("use_bvh", C.c_int32), ("precision", C.c_int32), ("pad", C.c_int32)]


PRECISION_F32, PRECISION_F64 = 0, 1
INTEGRATORS = {"path": 0, "direct": 1, "normals": 2}
TONE_MAPS = {"clamp": 0, "reinhard": 1, "aces": 2}

_VP = C.c_void_p


def _lib_path(name):
    return os.path.join(LIB_DIR, name)


class RtgError(RuntimeError):
    pass


_rtg = None


def load_rtg():
    """Loads librtg.so (the HIP back end); raises if it was not built."""
    global _rtg
    if _rtg is not None:
        return _rtg
    path = _lib_path("librtg.so")
    if not os.path.exists(path):
        raise RtgError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "rtg_open": (C.c_int32, [C.c_int, P(_VP)]),
        "rtg_close": (C.c_int32, [_VP]),
        "rtg_last_error": (C.c_char_p, [_VP]),
        "rtg_abi_version": (C.c_int, []),
        "rtg_device_info": (C.c_int32, [_VP, C.c_char_p, C.c_size_t, P(C.c_int), P(C.c_uint64)]),
        "rtg_scene_create": (C.c_int32, [_VP, P(SceneDesc), P(_VP)]),
        "rtg_scene_node_records": (C.c_int32, [P(SceneDesc), _VP, C.c_uint64, P(C.c_uint64)]),
        "rtg_scene_node_frame": (C.c_int32, [P(SceneDesc), P(C.c_float), P(C.c_int32)]),
        "rtg_scene_release": (C.c_int32, [_VP]),
        "rtg_scene_with_camera": (C.c_int32, [_VP, _VP, P(Camera), P(_VP)]),
        "rtg_scene_bytes": (C.c_uint64, [_VP]),
        "rtg_context_bytes": (C.c_uint64, [_VP]),
        "rtg_accum_create": (C.c_int32, [_VP, C.c_int32, C.c_int32, P(_VP)]),
        "rtg_accum_release": (C.c_int32, [_VP]),
        "rtg_accum_clear": (C.c_int32, [_VP, _VP]),
        "rtg_accum_read": (C.c_int32, [_VP, _VP, _VP, _VP]),
        "rtg_accum_min_samples": (C.c_uint32, [_VP, _VP]),
        "rtg_render_pass": (C.c_int32, [_VP, _VP, P(RenderSettings), C.c_uint32, C.c_uint32, _VP, _VP, P(Stats)]),
        "rtg_render_passes": (C.c_int32, [_VP, _VP, P(RenderSettings), P(C.c_uint32), C.c_uint32, _VP, _VP,
                                          P(C.c_uint32), P(Stats)]),
        "rtg_render_batch": (C.c_int32, [_VP, P(Job), C.c_uint32, P(Stats)]),
        "rtg_render_pass_static": (C.c_int32, [_VP, _VP, P(RenderSettings), C.c_uint32, C.c_uint32, _VP, _VP,
                                               _VP]),
        "rtg_resolve": (C.c_int32, [_VP, _VP, C.c_int32, C.c_double, _VP]),
        "rtg_buffer_create": (C.c_int32, [_VP, C.c_uint64, P(_VP)]),
        "rtg_buffer_release": (C.c_int32, [_VP]),
        "rtg_buffer_write": (C.c_int32, [_VP, _VP, _VP, C.c_uint64]),
        "rtg_buffer_read": (C.c_int32, [_VP, _VP, _VP, C.c_uint64]),
        "rtg_trace": (C.c_int32, [_VP, _VP, _VP, C.c_uint64, P(TraceOpts), _VP, P(Stats)]),
        "rtg_trace_host": (C.c_int32, [_VP, _VP, _VP, C.c_uint64, P(TraceOpts), _VP, P(Stats)]),
        "rtg_synchronize": (C.c_int32, [_VP]),
        "rtg_get_tuning": (C.c_int32, [_VP, P(Tuning)]),
        "rtg_set_tuning": (C.c_int32, [_VP, P(Tuning)]),
        "rtg_autotune": (C.c_int32, [_VP, P(Job), P(C.c_uint32), C.c_uint32, C.c_char_p, C.c_char_p,
                                     P(TuneResult)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _rtg = lib
    return lib


def rtg_symbols():
    """Names declared in include/rtg.h (checked by the CPU tests)."""
    return ["rtg_open", "rtg_close", "rtg_last_error", "rtg_abi_version", "rtg_device_info", "rtg_scene_create",
            "rtg_scene_node_records", "rtg_scene_node_frame",
            "rtg_scene_release", "rtg_scene_with_camera", "rtg_scene_bytes", "rtg_context_bytes", "rtg_accum_create", "rtg_accum_release", "rtg_accum_clear",
            "rtg_accum_read", "rtg_accum_min_samples", "rtg_render_pass", "rtg_render_passes", "rtg_render_batch",
            "rtg_render_pass_static", "rtg_resolve",
            "rtg_buffer_create",
            "rtg_buffer_release", "rtg_buffer_write", "rtg_buffer_read", "rtg_trace", "rtg_trace_host",
            "rtg_synchronize", "rtg_set_profiling", "rtg_set_profiling_kinds", "rtg_get_profile",
            "rtg_get_tuning", "rtg_set_tuning", "rtg_autotune"]


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def node_records(desc: "SceneDesc"):
    """The FP32 node records rtg_scene_create builds from `desc` (host only):
    an (n, 16) float32 array (rtg_scene_node_records)."""
    lib = load_rtg()
    n = C.c_uint64(0)
    st = lib.rtg_scene_node_records(C.byref(desc), None, 0, C.byref(n))
    if st != 0:
        raise RtgError(f"rtg status {st}: {lib.rtg_last_error(None).decode()}")
    out = np.zeros((n.value, 16), np.float32)
    if n.value:
        lib.rtg_scene_node_records(C.byref(desc), _ptr(out), n.value, C.byref(n))
    return out


def node_frame(desc: "SceneDesc"):
    """(origin float32[3], exp2) of the node records' frame: world
    coordinate = origin + record value * 2**exp2 (rtg_scene_node_frame)."""
    lib = load_rtg()
    o = (C.c_float * 3)()
    e = C.c_int32(0)
    st = lib.rtg_scene_node_frame(C.byref(desc), o, C.byref(e))
    if st != 0:
        raise RtgError(f"rtg status {st}: {lib.rtg_last_error(None).decode()}")
    return np.array(list(o), np.float32), e.value


class Context:
    """One HIP device + stream (rtg_open)."""

    def __init__(self, device=0):
        self.lib = load_rtg()
        h = C.c_void_p()
        self._check(self.lib.rtg_open(device, C.byref(h)), None)
        self.h = h

    def _check(self, st, h=-1):
        if st != 0:
            msg = self.lib.rtg_last_error(self.h if h == -1 else h)
            raise RtgError(f"rtg status {st}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.rtg_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        name = C.create_string_buffer(256)
        mp, mem = C.c_int(), C.c_uint64()
        self._check(self.lib.rtg_device_info(self.h, name, 256, C.byref(mp), C.byref(mem)))
        return {"name": name.value.decode(), "multiprocessors": mp.value, "hbm_bytes": mem.value}

    def scene(self, desc: SceneDesc):
        h = C.c_void_p()
        self._check(self.lib.rtg_scene_create(self.h, C.byref(desc), C.byref(h)))
        return DeviceScene(self, h)

    def accum(self, w, h):
        a = C.c_void_p()
        self._check(self.lib.rtg_accum_create(self.h, w, h, C.byref(a)))
        return Accum(self, a, w, h)

    def render_pass(self, scene, settings: RenderSettings, s0, s1, acc, cancel=None):
        st = Stats()
        cp = C.cast(C.pointer(cancel), C.c_void_p) if cancel is not None else None
        rc = self.lib.rtg_render_pass(self.h, scene.h, C.byref(settings), s0, s1, acc.h, cp, C.byref(st))
        if rc == -4:
            return st, True
        self._check(rc)
        return st, False

    def render_pass_static(self, scene, settings: RenderSettings, s0, s1, acc, cancel=None, stats=True):
        """rtg_render_pass_static: the pass with no device allocation (path
        state in the library's fixed device scratch). stats=False passes NULL
        (the call only enqueues when cancel is None too). Returns (Stats or
        None, cancelled)."""
        st = Stats() if stats else None
        cp = C.cast(C.pointer(cancel), C.c_void_p) if cancel is not None else None
        rc = self.lib.rtg_render_pass_static(self.h, scene.h, C.byref(settings), s0, s1, acc.h, cp,
                                             C.byref(st) if st is not None else None)
        if rc == -4:
            return st, True
        self._check(rc)
        return st, False

    def render_passes(self, scene, settings: RenderSettings, bounds, acc, cancel=None):
        """Passes [bounds[i], bounds[i + 1]) in one call (rtg_render_passes).
        Returns (total Stats, passes completed, cancelled)."""
        b = (C.c_uint32 * len(bounds))(*bounds)
        st = Stats()
        done = C.c_uint32(0)
        cp = C.cast(C.pointer(cancel), C.c_void_p) if cancel is not None else None
        rc = self.lib.rtg_render_passes(self.h, scene.h, C.byref(settings), b, len(bounds) - 1, acc.h, cp,
                                        C.byref(done), C.byref(st))
        if rc == -4:
            return st, done.value, True
        self._check(rc)
        return st, done.value, False

    def render_batch(self, jobs):
        """jobs: list of (scene, accum, settings, s0, s1[, p0, p1]). Returns
        one Stats per job (rtg_render_batch)."""
        arr = (Job * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            scene, acc, cfg, s0, s1 = j[:5]
            p0, p1 = (j[5], j[6]) if len(j) > 5 else (0, 0)
            arr[i] = Job(scene.h.value, acc.h.value, cfg, s0, s1, p0, p1)
        st = (Stats * max(1, len(jobs)))()
        self._check(self.lib.rtg_render_batch(self.h, arr, len(jobs), st))
        return list(st)[:len(jobs)]

    def device_bytes(self):
        """rtg_context_bytes: device memory the context holds now."""
        return self.lib.rtg_context_bytes(self.h)

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.rtg_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, chunk_paths=0, wide_paths=0, lanes=0):
        """Scheduling knobs (rtg_set_tuning); 0 = built-in default."""
        self._check(self.lib.rtg_set_tuning(self.h, C.byref(Tuning(chunk_paths, wide_paths, lanes, 0))))

    def autotune(self, calls, cache_path=None, key="default"):
        """rtg_autotune over a workload of batch calls: calls = list of job
        lists as render_batch takes them. Returns a dict with the chosen
        tuning, the default and tuned wall times and the trials timed."""
        flat = [j for call in calls for j in call]
        arr = (Job * max(1, len(flat)))()
        for i, j in enumerate(flat):
            scene, acc, cfg, s0, s1 = j[:5]
            p0, p1 = (j[5], j[6]) if len(j) > 5 else (0, 0)
            arr[i] = Job(scene.h.value, acc.h.value, cfg, s0, s1, p0, p1)
        sizes = (C.c_uint32 * len(calls))(*[len(c) for c in calls])
        r = TuneResult()
        self._check(self.lib.rtg_autotune(self.h, arr, sizes, len(calls),
                                          cache_path.encode() if cache_path else None, key.encode(), C.byref(r)))
        return {"tuning": r.best.as_dict(), "default_ms": r.default_ms, "best_ms": r.best_ms, "trials": r.trials}

    def trace(self, scene, rays: np.ndarray, mode=0, use_bvh=1, precision=PRECISION_F64):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 8)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        st = Stats()
        opts = TraceOpts(mode, use_bvh, precision, 0)
        self._check(self.lib.rtg_trace_host(self.h, scene.h, _ptr(rays), len(rays), C.byref(opts), _ptr(hits),
                                            C.byref(st)))
        return hits, st

    def synchronize(self):
        self._check(self.lib.rtg_synchronize(self.h))


class DeviceScene:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def nbytes(self):
        return self.ctx.lib.rtg_scene_bytes(self.h)

    def release(self):
        if self.h:
            self.ctx.lib.rtg_scene_release(self.h)
            self.h = None


class Accum:
    def __init__(self, ctx, h, w, hh):
        self.ctx, self.h, self.w, self.hgt = ctx, h, w, hh

    def read(self):
        s = np.zeros((self.hgt, self.w, 3), np.float64)
        c = np.zeros((self.hgt, self.w), np.uint32)
        self.ctx._check(self.ctx.lib.rtg_accum_read(self.ctx.h, self.h, _ptr(s), _ptr(c)))
        return s, c

    def clear(self):
        self.ctx._check(self.ctx.lib.rtg_accum_clear(self.ctx.h, self.h))

    def min_samples(self):
        return self.ctx.lib.rtg_accum_min_samples(self.ctx.h, self.h)

    def resolve(self, tone_map=0, exposure=1.0):
        out = np.zeros((self.hgt, self.w, 3), np.uint8)
        self.ctx._check(self.ctx.lib.rtg_resolve(self.ctx.h, self.h, tone_map, exposure, _ptr(out)))
        return out

    def release(self):
        if self.h:
            self.ctx.lib.rtg_accum_release(self.h)
            self.h = None


class HostArrays:
    """Owns numpy copies of a compiled scene and exposes them as a SceneDesc."""

    def __init__(self, prims, sampled, planes, materials, point_lights, lights, nodes, prim_indices, camera,
                 background):
        def arr(x, ctype):
            n = len(x)
            a = (ctype * max(n, 1))()
            for i, v in enumerate(x):
                a[i] = v
            return a, n

        self._keep = []
        d = SceneDesc()
        for field, ctype, data in (("prims", Primitive, prims), ("planes", Plane, planes),
                                   ("materials", Material, materials), ("point_lights", PointLight, point_lights),
                                   ("lights", LightRef, lights), ("nodes", BvhNode, nodes)):
            a, n = data if isinstance(data, tuple) else arr(data, ctype)
            self._keep.append(a)
            setattr(d, field, C.cast(a, C.POINTER(ctype)))
            setattr(d, "n_" + field, n)
        s = (C.c_uint8 * max(len(sampled), 1))(*sampled)
        pi = (C.c_uint32 * max(len(prim_indices), 1))(*prim_indices)
        self._keep += [s, pi]
        d.is_sampled_light = C.cast(s, C.POINTER(C.c_uint8))
        d.prim_indices = C.cast(pi, C.POINTER(C.c_uint32))
        d.n_prim_indices = len(prim_indices)
        d.camera = camera
        d.background = background
        self.desc = d


# ---------------------------------------------------------------------------
# Host library (include/rt_host.h, librthost.so)

class Overrides(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("rr_min_depth", C.c_int32), ("integrator", C.c_int32), ("light_sampling", C.c_int32),
                ("use_bvh", C.c_int32), ("seed", C.c_int64), ("exposure", C.c_double)]


class HostRenderStats(C.Structure):
    _fields_ = [("camera_samples", C.c_uint64), ("rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("rr_terminated", C.c_uint64), ("invalid_samples", C.c_uint64), ("node_visits", C.c_uint64),
                ("prim_tests", C.c_uint64), ("passes", C.c_uint64), ("elapsed_ms", C.c_double),
                ("device_ms", C.c_double), ("spp_done", C.c_uint32), ("cancelled", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQUEST_DTYPE = np.dtype([("scene", "<i4"), ("width", "<i4"), ("height", "<i4"), ("spp", "<i4"),
                          ("max_depth", "<i4"), ("max_pass", "<i4"), ("final_kind", "<i4"), ("pad", "<i4"),
                          ("seed", "<u8"), ("arrival_ms", "<f8")])
RESULT_DTYPE = np.dtype([("arrival_ms", "<f8"), ("start_ms", "<f8"), ("first_frame_ms", "<f8"), ("done_ms", "<f8"),
                         ("rays", "<u8"), ("shadow_rays", "<u8"), ("frames", "<u4"), ("spp_done", "<u4"),
                         ("image_hash", "<u8")])


class ServeStats(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("rays", C.c_uint64), ("ticks", C.c_uint32), ("max_active", C.c_uint32),
                ("peak_accum_bytes", C.c_uint64), ("gpu_ms", C.c_double), ("fixed_bytes", C.c_uint64),
                ("peak_footprint_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_rth = None


def load_host():
    """Loads librthost.so (which links librtg.so)."""
    global _rth
    if _rth is not None:
        return _rth
    path = _lib_path("librthost.so")
    if not os.path.exists(path):
        raise RtgError(f"{path} is missing: build it first")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "rt_validate": (C.c_char_p, [C.c_char_p]),
        "rt_message": (C.c_char_p, []),
        "rt_threads": (C.c_int, []),
        "rt_running": (C.c_int, []),
        "rt_cancel": (None, []),
        "rt_start": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32, _VP, _VP, _VP]),
        "rt_wait": (None, []),
        "rth_scene_load": (C.c_int, [C.c_char_p, P(Overrides), P(_VP)]),
        "rth_scene_free": (None, [_VP]),
        "rth_last_error": (C.c_char_p, []),
        "rth_scene_desc": (P(SceneDesc), [_VP]),
        "rth_scene_settings": (None, [_VP, C.c_int32, P(RenderSettings)]),
        "rth_scene_info": (None, [_VP, P(C.c_uint64)]),
        "rth_scene_bvh_build_ms": (C.c_double, [_VP]),
        "rth_dump_scene": (C.c_int64, [C.c_char_p, C.c_char_p, C.c_uint64]),
        "rth_render": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _VP, _VP, _VP,
                                 P(HostRenderStats)]),
        "rts_create": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint32, P(_VP)]),
        "rts_add_scene": (C.c_int, [_VP, _VP, P(C.c_int32)]),
        "rts_run": (C.c_int, [_VP, _VP, C.c_uint32, C.c_int32, _VP, P(ServeStats)]),
        "rts_set_budget": (C.c_int, [_VP, C.c_uint64]),
        "rts_free": (None, [_VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _rth = lib
    return lib


def host_symbols():
    return ["rt_validate", "rt_message", "rt_threads", "rt_running", "rt_cancel", "rt_start", "rt_wait",
            "rth_scene_load", "rth_scene_free", "rth_last_error", "rth_scene_desc", "rth_scene_settings",
            "rth_scene_info", "rth_scene_bvh_build_ms", "rth_dump_scene", "rth_render", "rts_create",
            "rts_add_scene", "rts_run", "rts_set_budget", "rts_free"]


class Server:
    """The served render loop (rts_*): scenes resident on one GPU, requests
    replayed in real time by the tick scheduler or by the plain loop."""

    def __init__(self, device=0, budget_bytes=0, tick_paths=0):
        self.lib = load_host()
        h = C.c_void_p()
        if self.lib.rts_create(device, budget_bytes, tick_paths, C.byref(h)):
            raise RtgError(self.lib.rth_last_error().decode())
        self.h = h
        self._scenes = []

    def add_scene(self, host_scene):
        i = C.c_int32()
        if self.lib.rts_add_scene(self.h, host_scene.h, C.byref(i)):
            raise RtgError(self.lib.rth_last_error().decode())
        self._scenes.append(host_scene)
        return i.value

    def run(self, requests, served=True):
        """requests: structured array of REQUEST_DTYPE sorted by arrival.
        Returns (results array of RESULT_DTYPE, stats dict)."""
        reqs = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
        res = np.zeros(len(reqs), RESULT_DTYPE)
        st = ServeStats()
        if self.lib.rts_run(self.h, _ptr(reqs), len(reqs), 1 if served else 0, _ptr(res), C.byref(st)):
            raise RtgError(self.lib.rth_last_error().decode())
        return res, st.as_dict()

    def set_budget(self, budget_bytes):
        """Footprint budget for the next run (0: 64 GiB)."""
        if self.lib.rts_set_budget(self.h, budget_bytes):
            raise RtgError(self.lib.rth_last_error().decode())

    def close(self):
        if self.h:
            self.lib.rts_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


FRAME_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.POINTER(C.c_uint8), C.c_void_p)
DONE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_void_p)


class HostScene:
    """A scene parsed, validated and compiled by the C++ host (Scene::compile
    + BVH build), viewable as an rtg SceneDesc."""

    def __init__(self, text, **ov):
        import json as _json
        lib = load_host()
        if not isinstance(text, str):
            text = _json.dumps(text)
        o = Overrides(-1, -1, -1, -1, -1, -1, -1, -1, -1, 0.0)
        for k, v in ov.items():
            if v is None:
                continue
            if k == "integrator":
                v = INTEGRATORS[v]
            if isinstance(v, bool):
                v = int(v)
            setattr(o, k, v)
        h = C.c_void_p()
        if lib.rth_scene_load(text.encode(), C.byref(o), C.byref(h)) != 0:
            raise RtgError(lib.rth_last_error().decode())
        self.lib, self.h = lib, h
        self.desc = lib.rth_scene_desc(h).contents
        self.desc._owner = self  # the arrays live in the host scene

    def settings(self, precision=PRECISION_F32):
        s = RenderSettings()
        self.lib.rth_scene_settings(self.h, precision, C.byref(s))
        return s

    def info(self):
        a = (C.c_uint64 * 10)()
        self.lib.rth_scene_info(self.h, a)
        keys = ["triangles", "spheres", "planes", "degenerate_skipped", "lights", "nodes", "leaves", "max_depth",
                "spp", "tone_map"]
        d = dict(zip(keys, list(a)))
        d["bvh_build_ms"] = self.lib.rth_scene_bvh_build_ms(self.h)
        return d

    def render(self, precision=PRECISION_F32, progressive=True, max_pass=32, device=0):
        cfg = self.settings(precision)
        s = np.zeros((cfg.height, cfg.width, 3))
        c = np.zeros((cfg.height, cfg.width), np.uint32)
        rgb = np.zeros((cfg.height, cfg.width, 3), np.uint8)
        st = HostRenderStats()
        if self.lib.rth_render(self.h, device, precision, 1 if progressive else 0, max_pass, _ptr(s), _ptr(c),
                               _ptr(rgb), C.byref(st)) != 0:
            raise RtgError(self.lib.rth_last_error().decode())
        return s, c, rgb, st.as_dict()

    def close(self):
        if self.h:
            self.lib.rth_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dump_scene(text):
    lib = load_host()
    buf = C.create_string_buffer(1 << 22)
    n = lib.rth_dump_scene(text.encode(), buf, len(buf))
    if n <= 0:
        raise RtgError(lib.rth_last_error().decode())
    return buf.value.decode()


def validate(text):
    return load_host().rt_validate(text.encode()).decode()


K_NAMES = ["generate", "extend", "shade", "shadow", "accumulate", "direct", "trace", "plan"]


class Profile(C.Structure):
    _fields_ = [("launches", C.c_uint64 * 8), ("ms", C.c_double * 8),
                ("closest_rays", C.c_uint64), ("closest_visits", C.c_uint64), ("closest_tests", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("shadow_visits", C.c_uint64), ("shadow_tests", C.c_uint64)]

    def as_dict(self):
        d = {k: {"launches": int(self.launches[i]), "ms": float(self.ms[i])} for i, k in enumerate(K_NAMES)}
        for k in ("closest_rays", "closest_visits", "closest_tests", "shadow_rays", "shadow_visits", "shadow_tests"):
            d[k] = int(getattr(self, k))
        return d


def set_profiling(ctx: "Context", on: bool, kinds=None):
    """Per-launch timing on/off; `kinds` (names of K_NAMES) limits it to
    those kernels (rtg_set_profiling_kinds)."""
    lib = ctx.lib
    if kinds is None:
        lib.rtg_set_profiling.restype, lib.rtg_set_profiling.argtypes = C.c_int32, [_VP, C.c_int32]
        ctx._check(lib.rtg_set_profiling(ctx.h, 1 if on else 0))
        return
    mask = 0
    for k in kinds:
        mask |= 1 << K_NAMES.index(k)
    lib.rtg_set_profiling_kinds.restype, lib.rtg_set_profiling_kinds.argtypes = C.c_int32, [_VP, C.c_uint32]
    ctx._check(lib.rtg_set_profiling_kinds(ctx.h, mask if on else 0))


def get_profile(ctx: "Context") -> Profile:
    lib = ctx.lib
    lib.rtg_get_profile.restype, lib.rtg_get_profile.argtypes = C.c_int32, [_VP, C.POINTER(Profile)]
    p = Profile()
    ctx._check(lib.rtg_get_profile(ctx.h, C.byref(p)))
    return p
