"""The batched workload: a seed expands to a list of render requests, each a
complete scene in the reference's JSON format.

Kinds (uniform): a Cornell box variant (wall colours, light radiance, block
rotations and camera field of view drawn per scene; 36 triangles, depth 8),
a torus with segments = 2 x segments_v and 5,000-200,000 triangles
(log-uniform; plane, sphere light and point light as in torus_c2.json,
depth 6), or a sphere field of 50-1,000 spheres (make_sphere_field.py with
a grid drawn per scene, depth 12). Resolution is drawn from 64^2, 128^2,
256^2, 512^2, 1024^2 and samples per pixel from 1, 4, 16, 64, 256; every
request has its own render seed. Draws come from a PCG32 stream, so a seed
always gives the same list.

    python make_batch.py [seed] [count]     # writes a JSON list to stdout
"""
import copy
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_sphere_field import Pcg32, build as sphere_field  # noqa: E402

RESOLUTIONS = (64, 128, 256, 512, 1024)
SPPS = (1, 4, 16, 64, 256)


def _load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def cornell(rng):
    j = _load("cornell_c1.json")
    m = j["materials"]
    m["red"]["albedo"] = [rng.uniform(0.4, 0.8), rng.uniform(0.02, 0.1), rng.uniform(0.02, 0.1)]
    m["green"]["albedo"] = [rng.uniform(0.05, 0.2), rng.uniform(0.3, 0.6), rng.uniform(0.05, 0.2)]
    m["light"]["intensity"] = rng.uniform(8.0, 20.0)
    for o in j["objects"]:
        if o["type"] == "box":
            o["rotate_y"] = rng.uniform(-30.0, 30.0)
    j["camera"]["vfov"] = rng.uniform(35.0, 45.0)
    j["name"] = "Cornell variant"
    return j


def torus(rng):
    j = _load("torus_c2.json")
    tris = math.exp(rng.uniform(math.log(5000.0), math.log(200000.0)))
    sv = max(3, int(round(math.sqrt(tris / 4.0))))
    mesh = next(o for o in j["objects"] if o["type"] == "mesh")
    mesh["segments"], mesh["segments_v"] = 2 * sv, sv
    j["materials"]["clay"]["albedo"] = [rng.uniform(0.3, 0.9) for _ in range(3)]
    light = next(o for o in j["objects"] if o.get("name") == "sphere light")
    light["center"] = [rng.uniform(-1.6, 1.6), rng.uniform(0.9, 1.6), rng.uniform(-1.2, 1.2)]
    j["name"] = f"Torus {2 * sv}x{sv}"
    return j


def spheres(rng):
    # a, b in [-half, half): about 0.95 * (2 half)^2 spheres survive the feature exclusion
    n = rng.uniform(50.0, 1000.0)
    half = max(4, int(round(math.sqrt(n / 0.95) / 2)))
    j = sphere_field(seed=rng.next_u32(), half=half)
    j["name"] = f"Sphere field {2 * half}x{2 * half}"
    return j


def build(seed=4, count=64):
    rng = Pcg32(seed)
    out = []
    for i in range(count):
        kind = rng.next_u32() % 3
        j = (cornell, torus, spheres)[kind](rng)
        res = RESOLUTIONS[rng.next_u32() % len(RESOLUTIONS)]
        spp = SPPS[rng.next_u32() % len(SPPS)]
        r = j["render"]
        r["width"] = r["height"] = res
        r["spp"] = spp
        r["seed"] = rng.next_u32() % 1000000000  # the scene reader takes integers up to 2e9
        out.append(copy.deepcopy(j))
    return out


if __name__ == "__main__":
    seed = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    count = int(sys.argv[2]) if len(sys.argv) > 2 else 64
    json.dump(build(seed, count), sys.stdout, indent=1)
    sys.stdout.write("\n")
