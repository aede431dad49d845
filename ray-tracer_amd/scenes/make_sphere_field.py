"""Generates sphere_field_c3.json: the depth-of-field sphere field workload.

22 x 22 grid, a, b in [-11, 11): centre (a + 0.9 U, r, b + 0.9 U) with
r ~ U(0.15, 0.25), dropped within 0.9 (in the ground plane) of the three
feature spheres; materials 70 % Lambertian (albedo U*U), 20 % metal (albedo
0.5 + 0.5 U, roughness U(0, 0.5)), 10 % dielectric (IOR 1.5). Feature
spheres: glass (0, 1, 0), diffuse (-4, 1, 0), mirror (4, 1, 0), radius 1.
One spherical emitter, a checker ground plane, thin-lens camera (aperture
0.1, focus 10) at (13, 2, 3) looking at the origin; 1280x720, 128 spp,
depth 12, seed 3. Draws come from a PCG32 stream seeded with 3, so the file
is reproducible: python make_sphere_field.py > sphere_field_c3.json
"""
import json
import sys


class Pcg32:
    """PCG32 XSH-RR (O'Neill 2014), 64-bit state, 32-bit output."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed, stream=0xDA3E39CB94B95BDB):
        self.inc = ((stream << 1) | 1) & self.MASK
        self.state = 0
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * (self.next_u32() / 4294967296.0)


def build(seed=3, half=11):
    """The scene as a dict; `half` sets the grid, a, b in [-half, half)."""
    rng = Pcg32(seed)
    mats = {
        "ground": {"type": "lambertian", "albedo": [0.5, 0.5, 0.5], "albedo2": [0.32, 0.36, 0.3],
                   "checker_scale": 0.5},
        "glass": {"type": "dielectric", "ior": 1.5},
        "big_diffuse": {"type": "lambertian", "albedo": [0.4, 0.2, 0.1]},
        "big_metal": {"type": "metal", "albedo": [0.7, 0.6, 0.5], "roughness": 0.0},
        "lamp": {"type": "emissive", "color": [1.0, 0.92, 0.8], "intensity": 6.0},
    }
    objs = [{"type": "plane", "name": "ground", "point": [0, 0, 0], "normal": [0, 1, 0], "material": "ground"}]
    features = [(0.0, 0.0), (-4.0, 0.0), (4.0, 0.0)]
    k = 0
    for a in range(-half, half):
        for b in range(-half, half):
            choose = rng.uniform()
            x = a + 0.9 * rng.uniform()
            z = b + 0.9 * rng.uniform()
            r = rng.uniform(0.15, 0.25)
            if any((x - fx) ** 2 + (z - fz) ** 2 <= 0.9 ** 2 for fx, fz in features):
                continue
            name = f"m{k}"
            if choose < 0.7:
                mats[name] = {"type": "lambertian", "albedo": [rng.uniform() * rng.uniform() for _ in range(3)]}
            elif choose < 0.9:
                mats[name] = {"type": "metal", "albedo": [0.5 + 0.5 * rng.uniform() for _ in range(3)],
                              "roughness": rng.uniform(0.0, 0.5)}
            else:
                name = "glass"
            objs.append({"type": "sphere", "name": f"sphere {k}", "center": [x, r, z], "radius": r, "material": name})
            k += 1
    objs += [
        {"type": "sphere", "name": "big glass", "center": [0, 1, 0], "radius": 1.0, "material": "glass"},
        {"type": "sphere", "name": "big diffuse", "center": [-4, 1, 0], "radius": 1.0, "material": "big_diffuse"},
        {"type": "sphere", "name": "big metal", "center": [4, 1, 0], "radius": 1.0, "material": "big_metal"},
        {"type": "sphere", "name": "lamp", "center": [3, 7, 5], "radius": 1.2, "material": "lamp"},
    ]
    return {
        "name": "Sphere field (depth of field)",
        "description": "~480 spheres radius U(0.15,0.25) on a jittered 22x22 grid, 70/20/10 Lambertian/metal/"
                       "dielectric, three feature spheres, one spherical light; thin lens aperture 0.1 focus 10; "
                       "1280x720, 128 spp, depth 12, seed 3.",
        "camera": {"look_from": [13, 2, 3], "look_at": [0, 0, 0], "vfov": 20, "aperture": 0.1, "focus_dist": 10},
        "render": {"width": 1280, "height": 720, "spp": 128, "max_depth": 12, "rr_min_depth": 3,
                   "integrator": "path", "light_sampling": True, "seed": 3, "tone_map": "aces"},
        "background": {"type": "gradient", "top": [0.5, 0.7, 1.0], "bottom": [1, 1, 1], "intensity": 0.6},
        "materials": mats,
        "objects": objs,
    }


if __name__ == "__main__":
    json.dump(build(int(sys.argv[1]) if len(sys.argv) > 1 else 3), sys.stdout, indent=1)
    sys.stdout.write("\n")
