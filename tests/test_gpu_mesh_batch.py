"""GPU parity on triangle meshes, marched primitives and multi-scene batches.

Meshes come in through the binary triangle API (the PLY path builds the
same records, a CPU test checks that); marched primitives (torus, superellipsoid) are
checked in both intersect modes: the reference's SDF march (FP64 kernel vs
oracle march: t within 1e-9 relative) and the exact mode (GPU closed-form
quartic / bracketed roots vs the oracle's independent dense-bracketing
root finder: t within 1e-7 relative).
"""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi, workloads
import oracle
from test_gpu_parity import compare_hits, random_rays

pytestmark = pytest.mark.gpu


def mesh_world(freq=8):
    toml, add_mesh, info = workloads.batch_scene(3, seed=5)
    cfg, cam, world = lm.load_scene(toml)
    v, f = workloads.displaced_sphere(freq, [0.0, 1.2, 0.0], 1.0, 77)
    m = world.add_material(workloads._microfacet([0.7, 0.6, 0.5], 0.3))
    world.add_triangles(v, f, material=m)
    return cfg, cam, world


@pytest.mark.parametrize("path", ["wavefront", "plain64"])
def test_mesh_ray_batch(gpu, path):
    cfg, cam, world = mesh_world(freq=12)
    rays = random_rays(40000, [-2.5, 0.05, -2.5], [2.5, 3.0, 3.0], seed=11)
    ho = oracle.trace(world, rays)
    hg = lm.trace_rays(world, rays, path=path)
    r = compare_hits(rays, ho, hg, fp32=(path == "wavefront"))
    print(path, r)
    assert r["hits"] > 10000
    assert r["bad_t"] == 0
    # shared mesh edges: barycentric bounds are inclusive on both triangles
    assert r["flips"] <= (0 if path == "plain64" else int(5e-4 * r["n"]))


def test_mesh_render_parity(gpu):
    cfg, cam, world = mesh_world(freq=10)
    cfg.width, cfg.height, cfg.samples_per_pixel = 48, 48, 9
    ref, var, ost = oracle.render(world, cfg)
    r64 = lm.render(cfg, cam, world, path="plain64")
    assert np.linalg.norm(r64.rows - ref) / np.linalg.norm(ref) < 1e-6
    assert r64.stats["shadow_rays"] == ost["shadow_rays"]
    r32 = lm.render(cfg, cam, world, path="wavefront")
    frac = float((np.abs(r32.rows - ref) <= 3.0 * np.sqrt(2.0 * var / 9) + 1e-6).mean())
    assert frac >= 0.99


MARCHED = """
[render]
width = 32
height = 32
samples = 4
max_depth = 6
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
[[torus]]
center = [1.2, 0.3, 0.4]
major_radius = 0.6
minor_radius = 0.2
material = { type = "metal", color = [0.8, 0.8, 0.8], fuzz = 0.1 }
[[superellipsoid]]
center = [0.3, 0.5, -0.5]
scale = [0.4, 0.5, 0.4]
e1 = 0.5
e2 = 1.5
material = { type = "pbr", color = [0.3, 0.6, 0.8], roughness = 0.3 }
[[light]]
position = [0.0, 3.0, 2.0]
radius = 0.3
intensity = 30.0
"""


@pytest.mark.parametrize("mode", [_abi.ISECT_REFERENCE, _abi.ISECT_EXACT])
def test_marched_primitives_fp64(gpu, mode):
    cfg, cam, world = lm.load_scene(MARCHED)
    rng = np.random.default_rng(3)
    n = 20000
    o = rng.uniform([-3, 0.2, 3], [3, 1.5, 5], size=(n, 3))
    tgt = rng.uniform([-1.8, 0.0, -1.0], [1.8, 0.9, 1.0], size=(n, 3))
    rays = np.concatenate([o, tgt - o, np.full((n, 1), 0.001), np.full((n, 1), np.inf)], axis=1)
    exact = mode == _abi.ISECT_EXACT
    ho = oracle.trace(world, rays, exact=exact)
    hg = lm.trace_rays(world, rays, path="plain64", intersect_mode=mode)
    same = ho["prim"] == hg["prim"]
    hit = same & (ho["prim"] >= 0)
    rel = np.abs(hg["t"][hit] - ho["t"][hit]) / np.abs(ho["t"][hit])
    tol = 1e-7 if exact else 1e-9
    print(mode, "mismatch", (~same).sum(), "max rel", rel.max())
    # marched hits: the march stops within its own tolerance; FP64 GPU and CPU
    # take the same steps, so they agree far inside it
    assert (~same).sum() <= int(2e-4 * n)
    assert (rel <= tol).mean() >= 0.999


def test_marched_render_statistical(gpu):
    cfg, cam, world = lm.load_scene(MARCHED)
    cfg.samples_per_pixel = 16
    ref, var, _ = oracle.render(world, cfg, exact=True)
    r = lm.render(cfg, cam, world, path="wavefront", intersect_mode=_abi.ISECT_EXACT)
    frac = float((np.abs(r.rows - ref) <= 3.0 * np.sqrt(2.0 * var / 16) + 1e-6).mean())
    assert frac >= 0.98
    r64 = lm.render(cfg, cam, world, path="plain64", intersect_mode=_abi.ISECT_EXACT)
    assert np.linalg.norm(r64.rows - ref) / np.linalg.norm(ref) < 1e-4


def test_batch_matches_individual_renders(gpu):
    cfgs, worlds = [], []
    for i in range(4):
        toml, add_mesh, info = workloads.batch_scene(i, seed=9, max_res=64)
        cfg, cam, world = lm.load_scene(toml)
        v, f = workloads.displaced_sphere(6 + i, [0.0, 1.2, 0.0], 1.0, 100 + i)
        world.add_triangles(v, f, material=world.add_material(workloads._microfacet([0.6, 0.6, 0.6], 0.4)))
        cfg.samples_per_pixel = 4
        cfgs.append(cfg)
        worlds.append(world)
    outs, st = lm.render_batch(cfgs, worlds)
    for i, (c, w) in enumerate(zip(cfgs, worlds)):
        single = lm.render(c, None, w, scene_key=i)
        assert np.array_equal(outs[i], single.rows), i
    assert st["paths"] == sum(c.width * c.height * 4 for c in cfgs)


def test_marched_wavefront_equals_plain_fp32(gpu):
    # the wavefront's compacted marched pass (candidate queues, one kind per
    # launch, closest hit by atomic min) against the plain FP32 kernel's
    # in-place marched traversal: the same FP32 hits per path
    cfg, cam, world = lm.load_scene(MARCHED)
    cfg.samples_per_pixel = 16
    a = lm.render(cfg, cam, world, path="wavefront", intersect_mode=_abi.ISECT_EXACT)
    b = lm.render(cfg, cam, world, path="plain32", intersect_mode=_abi.ISECT_EXACT)
    np.testing.assert_allclose(a.rows, b.rows, rtol=1e-6, atol=1e-7)
    assert a.stats["extension_rays"] == b.stats["extension_rays"]
    assert a.stats["shadow_rays"] == b.stats["shadow_rays"]


def _torus_row(n):
    # n tori one behind the other along the camera's view: camera rays cross
    # up to n torus boxes, more than a ray may queue
    body = MARCHED.split("[[torus]]")[0]
    for k in range(n):
        body += ("[[torus]]\n" f"center = [0.0, 0.3, {1.0 - 0.15 * k}]\n" "major_radius = 0.5\nminor_radius = 0.15\n"
                 'material = { type = "lambertian", color = [0.7, 0.5, 0.3] }\n')
    body += '[[light]]\nposition = [0.0, 3.0, 2.0]\nradius = 0.3\nintensity = 30.0\n'
    return body


@pytest.mark.parametrize("pool", [0, 4096, 4000])
def test_marched_overflow_fallback(gpu, pool):
    # rays with more candidates than the per-ray limit, and (small pool)
    # candidates beyond the queue capacity, go to the fallback kernel; the
    # image must still be the plain FP32 kernel's.  A pool that is not a
    # multiple of the workgroup size leaves lanes past the pool in the last
    # workgroup: they must stay off the fallback list.
    cfg, cam, world = lm.load_scene(_torus_row(16))
    cfg.samples_per_pixel = 4
    a = lm.render(cfg, cam, world, path="wavefront", intersect_mode=_abi.ISECT_EXACT, pool_paths=pool)
    b = lm.render(cfg, cam, world, path="plain32", intersect_mode=_abi.ISECT_EXACT)
    np.testing.assert_allclose(a.rows, b.rows, rtol=1e-6, atol=1e-7)
    assert a.stats["extension_rays"] == b.stats["extension_rays"]


def test_marched_wavefront_equals_plain_fp32_reference_mode(gpu):
    # the same in the reference's march semantics (SDF marches per kind)
    cfg, cam, world = lm.load_scene(MARCHED)
    cfg.samples_per_pixel = 16
    a = lm.render(cfg, cam, world, path="wavefront", intersect_mode=_abi.ISECT_REFERENCE)
    b = lm.render(cfg, cam, world, path="plain32", intersect_mode=_abi.ISECT_REFERENCE)
    np.testing.assert_allclose(a.rows, b.rows, rtol=1e-6, atol=1e-7)
    assert a.stats["extension_rays"] == b.stats["extension_rays"]
