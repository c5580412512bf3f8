"""Seeded synthetic workloads, emitted as scene TOML (plus ASCII PLY meshes)
that the reference's loader accepts, so the same text drives the GPU path
and the CPU oracle.

All randomness comes from `SeededRng` (SplitMix64 sequence, Steele et al.
2014) keyed by (master seed, workload, scene, object), so every workload is
reproducible from its name and seed.
"""
import math
import os

MASK64 = (1 << 64) - 1


def _mix(z):
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & MASK64
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB & MASK64
    return z ^ (z >> 31)


class SeededRng:
    """SplitMix64 stream; uniform() returns [0, 1) doubles (53 bits)."""

    def __init__(self, *key):
        s = 0x9E3779B97F4A7C15
        for k in key:
            s = _mix((s ^ (int(k) & MASK64)) + 0x632BE59BD9B4E019 & MASK64)
        self.state = s

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        return _mix(self.state)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * 2.0 ** -53)

    def randint(self, n):
        return self.next_u64() % n


def _v(a):
    return "[" + ", ".join(repr(float(x)) for x in a) + "]"


def _header(width, height, spp, depth, seed, background, camera):
    lines = ["[render]", f"width = {width}", f"height = {height}", f"samples = {spp}",
             f"max_depth = {depth}", f"seed = {seed}", f"background = {background}", "", "[camera]"]
    for k, v in camera.items():
        lines.append(f"{k} = {_v(v) if isinstance(v, (list, tuple)) else repr(float(v))}")
    lines.append("")
    return lines


def _mat(m):
    parts = [f'type = "{m["type"]}"']
    for k, v in m.items():
        if k == "type":
            continue
        parts.append(f"{k} = {_v(v) if isinstance(v, (list, tuple)) else repr(float(v))}")
    return "{ " + ", ".join(parts) + " }"


def cornell_box(width=256, height=256, spp=64, depth=8, seed=42):
    """Cornell-style box: five axis-aligned quad walls of a 555-unit cube,
    a 130x105 emissive quad light (15,15,15) under the ceiling, two white
    Lambertian boxes 165^3 (rotated 18 deg) and 165x330x165 (rotated -15 deg),
    black background, camera (278,278,-800) -> (278,278,0), fov 40."""
    white = {"type": "lambertian", "color": [0.73, 0.73, 0.73]}
    red = {"type": "lambertian", "color": [0.65, 0.05, 0.05]}
    green = {"type": "lambertian", "color": [0.12, 0.45, 0.15]}
    L = _header(width, height, spp, depth, seed, '{ type = "black" }',
                {"look_from": [278, 278, -800], "look_at": [278, 278, 0], "vfov": 40, "aperture": 0,
                 "focus_dist": 1})
    quads = [
        ([0, 0, 0], [555, 0, 0], [0, 0, 555], white),      # floor
        ([0, 555, 0], [555, 0, 0], [0, 0, 555], white),    # ceiling
        ([0, 0, 555], [555, 0, 0], [0, 555, 0], white),    # back
        ([0, 0, 0], [0, 555, 0], [0, 0, 555], red),        # x = 0 wall
        ([555, 0, 0], [0, 555, 0], [0, 0, 555], green),    # x = 555 wall
        ([213, 554, 227], [130, 0, 0], [0, 0, 105], {"type": "emissive", "color": [15, 15, 15], "intensity": 1}),
    ]
    for q, u, v, m in quads:
        L += ["[[quad]]", f"q = {_v(q)}", f"u = {_v(u)}", f"v = {_v(v)}", f"material = {_mat(m)}", ""]
    L += ["[[box]]", "min = [0.0, 0.0, 0.0]", "max = [165.0, 165.0, 165.0]", f"material = {_mat(white)}",
          "rotate_y = 18.0", "translate = [130.0, 0.0, 65.0]", ""]
    L += ["[[box]]", "min = [0.0, 0.0, 0.0]", "max = [165.0, 330.0, 165.0]", f"material = {_mat(white)}",
          "rotate_y = -15.0", "translate = [265.0, 0.0, 295.0]", ""]
    return "\n".join(L)


def random_spheres(width=1920, height=1080, spp=128, depth=50, seed=42, grid=11):
    """Random-sphere scene: ground sphere r=1000 (albedo 0.5), a grid of
    r=0.2 spheres at integer x,z in [-grid, grid) with U[0,0.9] jitter
    (80% Lambertian albedo U^2, 15% metal U[.5,1] fuzz U[0,.5], 5% glass
    1.5), glass / Lambertian (0.4,0.2,0.1) / metal (0.7,0.6,0.5) r=1 spheres,
    sky gradient, camera (13,2,3) -> origin, fov 20, aperture 0.1, focus 10."""
    rng = SeededRng(seed, 1)
    L = _header(width, height, spp, depth, seed, '{ type = "sky" }',
                {"look_from": [13, 2, 3], "look_at": [0, 0, 0], "vfov": 20, "aperture": 0.1, "focus_dist": 10})

    def sphere(c, r, m):
        L.extend(["[[sphere]]", f"center = {_v(c)}", f"radius = {float(r)!r}", f"material = {_mat(m)}", ""])

    sphere([0, -1000, 0], 1000, {"type": "lambertian", "color": [0.5, 0.5, 0.5]})
    sphere([0, 1, 0], 1, {"type": "dielectric", "refraction_index": 1.5})
    sphere([-4, 1, 0], 1, {"type": "lambertian", "color": [0.4, 0.2, 0.1]})
    sphere([4, 1, 0], 1, {"type": "metal", "color": [0.7, 0.6, 0.5], "fuzz": 0.0})
    for a in range(-grid, grid):
        for b in range(-grid, grid):
            choose = rng.uniform()
            c = [a + 0.9 * rng.uniform(), 0.2, b + 0.9 * rng.uniform()]
            if math.dist(c, [4, 0.2, 0]) <= 0.9:
                continue
            if choose < 0.8:
                alb = [rng.uniform() * rng.uniform() for _ in range(3)]
                m = {"type": "lambertian", "color": alb}
            elif choose < 0.95:
                m = {"type": "metal", "color": [rng.uniform(0.5, 1.0) for _ in range(3)],
                     "fuzz": rng.uniform(0, 0.5)}
            else:
                m = {"type": "dielectric", "refraction_index": 1.5}
            sphere(c, 0.2, m)
    return "\n".join(L)


WORKLOADS = {
    "cornell": cornell_box,
    "random_spheres": random_spheres,
}
