"""Writes tests/golden/highspp.npz: high-sample-count image estimates of the
workload scenes rendered by the REFERENCE renderer itself
(oracle/_ref/librtref.so, compiled from the reference's sources by
oracle/Makefile), for the statistical image-parity tests of the FP32 device
path. Run in the container that has the reference checkout:

    make -C oracle && python tests/golden/make_highspp.py

For each case it stores, per pixel and channel, the mean of the case's
samples (16384, or 65536 for the Cornell box, whose next-event estimates
near the ceiling light are heavy-tailed, so fewer samples underestimate
their variance) and the unbiased sample variance of one sample's radiance (from the sum of
squares kept beside the reference's sum; render_tile's loop,
src/rendering/renderer.cpp:21-56, through the reference's radiance(),
src/rendering/integrator.cpp:184-198). Data only: inputs are the committed
scene files plus the overrides listed in the meta record.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import orc  # noqa: E402
import scenes  # noqa: E402

OUT = os.path.join(HERE, "highspp.npz")
# (case, scene file, width, height, integrator, samples per pixel)
CASES = [
    ("c1_path", "cornell_c1.json", 64, 64, "path", 65536),
    ("c2_path", "torus_c2.json", 64, 64, "path", 16384),
    ("c3_path", "sphere_field_c3.json", 96, 54, "path", 16384),
    ("c1_direct", "cornell_c1.json", 64, 64, "direct", 65536),
    ("c3_direct", "sphere_field_c3.json", 96, 54, "direct", 16384),
    ("c2_normals", "torus_c2.json", 64, 64, "normals", 16384),
]


def main():
    assert orc.have_ref(), "oracle/_ref/librtref.so missing: run make -C oracle in the reference container"
    out, meta = {}, {}
    threads = len(os.sched_getaffinity(0))
    for case, name, w, h, integ, spp in CASES:
        r = orc.RefScene(json.dumps(scenes.load(name)), width=w, height=h, integrator=integ)
        t0 = time.time()
        s, q, c, st = r.render_moments(0, spp, w, h, threads=threads)
        dt = time.time() - t0
        assert (c == spp).all()
        mean = s / spp
        var = np.maximum(q - s * mean, 0.0) / (spp - 1)
        out[case + "_mean"] = mean.astype(np.float32)
        out[case + "_var"] = var.astype(np.float32)
        meta[case] = {"scene": name, "width": w, "height": h, "integrator": integ, "spp": spp,
                      "seed": int(scenes.load(name).get("render", {}).get("seed", 1)),
                      "channel_mean": [float(x) for x in mean.reshape(-1, 3).mean(0)],
                      "stats": {"camera_samples": int(st[0]), "rays": int(st[1]), "shadow_rays": int(st[2]),
                                "rr_terminated": int(st[3]), "invalid_samples": int(st[4])},
                      "seconds": round(dt, 2), "threads": threads}
        print(case, meta[case], flush=True)
    out["meta"] = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
