"""Asynchronous renders on the device: RenderSession / GpuRenderer through
the native check program (the reference's tests/test_render.cpp:211-266:
progressive counts, cancel within 1 s of a 100000-spp render and restart,
RAII join, and a cancel raised while the whole schedule runs in one device
call), and the WebAssembly-shaped C ABI rt_start / rt_cancel / rt_wait /
rt_running / rt_message (apps/rt_wasm.cpp:114-184) through ctypes."""
import ctypes as C
import json
import os
import subprocess
import threading
import time

import numpy as np
import pytest

import scenes
from _pkg import rtg

pytestmark = pytest.mark.gpu
CHECK = os.path.join(scenes.ROOT, "tests", "native", "session_check")


def test_native_session_checks():
    assert os.path.exists(CHECK), "tests/native/session_check is not built (python -c 'import __graft_entry__ as g; g.build()')"
    r = subprocess.run([CHECK, scenes.ASSET_DIR], capture_output=True, timeout=120)
    out = r.stdout.decode()
    print(out, r.stderr.decode()[-2000:])
    assert r.returncode == 0, out
    for name in ("progressive_counts", "cancel_within_1s", "cancelled_buffer_consistent", "restart_completes",
                 "destructor_cancels_and_joins", "cancel_one_call"):
        assert f"ok {name}" in out


class Recorder:
    """Collects rt_start's frame and done callbacks (called on the render
    thread)."""

    def __init__(self):
        self.frames, self.done, self.lock = [], [], threading.Lock()
        self.frame_fn = rtg.FRAME_FN(self._frame)
        self.done_fn = rtg.DONE_FN(self._done)

    def _frame(self, user, job, rgba, info):
        fi = C.cast(info, C.POINTER(FrameInfo)).contents
        n = fi.width * fi.height * 4
        img = np.ctypeslib.as_array(rgba, shape=(n,)).copy().reshape(fi.height, fi.width, 4)
        with self.lock:
            self.frames.append((job, fi.pass_, fi.spp_done, bool(fi.finished), bool(fi.cancelled), img))

    def _done(self, user, job, info):
        di = C.cast(info, C.POINTER(DoneInfo)).contents
        with self.lock:
            self.done.append((job, bool(di.cancelled), di.spp_done, di.rays, di.shadow_rays))


class FrameInfo(C.Structure):
    _fields_ = [("pass_", C.c_int32), ("spp_done", C.c_uint32), ("spp_target", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("finished", C.c_int32), ("cancelled", C.c_int32), ("pad", C.c_int32),
                ("elapsed_ms", C.c_double), ("rays", C.c_uint64)]


class DoneInfo(C.Structure):
    _fields_ = [("cancelled", C.c_int32), ("spp_done", C.c_uint32), ("elapsed_ms", C.c_double),
                ("rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("invalid_samples", C.c_uint64),
                ("mrays_per_sec", C.c_double), ("threads", C.c_int32), ("pad", C.c_int32)]


def scene_text(w, h, spp):
    j = scenes.load_asset("cornell_box.json")
    j["render"].update({"width": w, "height": h, "spp": spp})
    return json.dumps(j)


def test_rt_start_cancel_wait():
    lib = rtg.load_host()
    rec = Recorder()
    assert lib.rt_threads() == 1
    assert lib.rt_validate(b'{"render": {"width": 0}}').decode() == "render.width: must be in [1, 8192]"
    assert lib.rt_validate(scene_text(8, 8, 1).encode()).decode() == ""

    assert lib.rt_start(7, scene_text(128, 96, 100000).encode(), 0, rec.frame_fn, rec.done_fn, None) == 0
    time.sleep(0.15)
    assert lib.rt_running() == 1
    assert lib.rt_start(8, scene_text(8, 8, 1).encode(), 0, rec.frame_fn, rec.done_fn, None) == 1
    assert lib.rt_message().decode() == "a render is still running"
    t0 = time.perf_counter()
    lib.rt_cancel()
    lib.rt_wait()
    ms = (time.perf_counter() - t0) * 1e3
    print(f"cancel -> wait returned in {ms:.1f} ms after {len(rec.frames)} frames")
    assert ms < 1000 and lib.rt_running() == 0
    assert len(rec.done) == 1 and rec.done[0][0] == 7 and rec.done[0][1] is True
    last = rec.frames[-1]
    assert last[4] is True and last[3] is False and last[2] == rec.done[0][2]
    assert all(f[0] == 7 for f in rec.frames) and np.all(last[5][..., 3] == 255)
    spp = [f[2] for f in rec.frames[:-1]]
    assert spp == sorted(spp) and spp[:3] == [1, 2, 4]

    # A small job after the cancelled one runs to completion, and its last
    # frame is the image GpuRenderer renders for the same scene.
    rec2 = Recorder()
    text = scene_text(48, 32, 4)
    assert lib.rt_start(9, text.encode(), 0, rec2.frame_fn, rec2.done_fn, None) == 0
    lib.rt_wait()
    assert [f[2] for f in rec2.frames] == [1, 2, 4] and rec2.frames[-1][3] is True
    assert rec2.done == [(9, False, 4, rec2.done[0][3], rec2.done[0][4])] and rec2.done[0][3] > 0
    hs = rtg.HostScene(text)
    rgb = hs.render(progressive=True)[2]
    assert np.array_equal(rec2.frames[-1][5][..., :3], rgb)
