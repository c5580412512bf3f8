"""Writes the golden ray-batch fixtures under tests/golden/ from the CPU FP64
oracle: for each scene, seeded rays and the closest hit (t, descriptor prim
id, u, v, normal) in both intersect modes -- the reference's marchers and the
exact roots.  Each fixture carries its scene's TOML text, so it is self-
contained.  Run from the repository root: python tests/golden/make_golden.py
[name ...] (all fixtures, or the named ones).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import luminara_amd as lm  # noqa: E402
from luminara_amd import workloads  # noqa: E402
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

MARCHED = """
[render]
width = 32
height = 32
background = { type = "sky" }
[camera]
look_from = [0.0, 2.0, 6.0]
look_at = [0.0, 0.3, 0.0]
vfov = 40.0
[[plane]]
point = [0.0, 0.0, 0.0]
normal = [0.0, 1.0, 0.0]
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
[[torus]]
center = [-1.0, 0.3, 0.0]
major_radius = 0.5
minor_radius = 0.15
material = { type = "lambertian", color = [0.8, 0.3, 0.3] }
[[superellipsoid]]
center = [0.3, 0.5, -0.5]
scale = [0.4, 0.5, 0.4]
e1 = 0.5
e2 = 1.5
material = { type = "pbr", color = [0.3, 0.6, 0.8], roughness = 0.3 }
[[superellipsoid]]
center = [1.4, 0.45, 0.3]
scale = [0.35, 0.45, 0.35]
e1 = 2.5
e2 = 2.5
material = { type = "lambertian", color = [0.6, 0.6, 0.3] }
[[spring]]
center = [-0.2, 0.0, 1.0]
coil_radius = 0.35
tube_radius = 0.06
pitch = 0.25
turns = 4.0
material = { type = "metal", color = [0.8, 0.8, 0.8], fuzz = 0.0 }
"""


def rays_toward(rng, n, lo_o, hi_o, lo_t, hi_t, unit=False):
    o = rng.uniform(lo_o, hi_o, size=(n, 3))
    d = rng.uniform(lo_t, hi_t, size=(n, 3)) - o
    if unit:
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d, np.full((n, 1), 0.001), np.full((n, 1), np.inf)], axis=1)


def scenes():
    rng = np.random.default_rng(2024)
    yield "cornell", workloads.cornell_box(64, 64, 4, 8, 42), rays_toward(
        rng, 1500, [100, 100, -800], [455, 455, -300], [0, 0, 0], [555, 555, 555])
    yield "random_spheres", workloads.random_spheres(64, 36, 4, 8, 42), rays_toward(
        rng, 1500, [-13, 0.5, -4], [13, 3, 4], [-11, -0.5, -11], [11, 1.5, 11])
    # unit directions: the reference's marchers step t by the SDF value
    yield "marched", MARCHED, rays_toward(rng, 1500, [-3, 0.2, 3], [3, 1.5, 5], [-1.8, 0.0, -1.0], [1.8, 1.0, 1.5],
                                          unit=True)
    # the same scene with every ray component (origin, direction, t_min)
    # exactly representable in FP32: the FP32 kernels then trace the very
    # rays the oracle traced, which the reference's marchers need (they are
    # chaotic in their input: a ray rounded to FP32 takes other steps)
    rays = rays_toward(np.random.default_rng(2025), 1500, [-3, 0.2, 3], [3, 1.5, 5], [-1.8, 0.0, -1.0],
                       [1.8, 1.0, 1.5], unit=True)
    rays[:, :7] = rays[:, :7].astype(np.float32).astype(np.float64)
    yield "marched_f32", MARCHED, rays


def main(only=None):
    for name, toml, rays in scenes():
        if only and name not in only:
            continue
        _, _, world = lm.load_scene(toml)
        out = {"toml": np.array(toml), "rays": rays}
        for mode, exact in (("reference", False), ("exact", True)):
            h = oracle.trace(world, rays, exact=exact)
            out[f"{mode}_t"] = h["t"]
            out[f"{mode}_prim"] = h["prim"]
            out[f"{mode}_uv"] = np.stack([h["u"], h["v"]], 1)
            out[f"{mode}_normal"] = h["normal"]
        path = os.path.join(HERE, f"rays_{name}.npz")
        np.savez_compressed(path, **out)
        print(path, len(rays), "rays,", int((out["exact_prim"] >= 0).sum()), "hits")


if __name__ == "__main__":
    main(sys.argv[1:])  # optional fixture names: write only those
