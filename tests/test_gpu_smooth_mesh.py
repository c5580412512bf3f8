"""Smooth-shaded meshes on the GPU (MeshTriangle with vertex normals and
texture coordinates, src/obj.rs:250-313): a textured geodesic sphere with
radial corner normals and spherical uvs, and a GGX displaced sphere with
corner normals only.

Bars: ray batches against the oracle (plain FP64 kernel: normals to 1e-12,
uvs to 1e-9; FP32 wavefront: the FP32 hit tolerances of the other ray-batch
tests, normals to 1e-5, uvs to 1e-4); images as in the other image parity
tests (plain FP64 vs oracle relRMSE < 1e-6 with equal ray counts, wavefront
within 3 standard errors on >= 99 % of the pixel channels, wavefront equal to
the plain FP32 kernel)."""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import workloads
import oracle
from test_gpu_parity import random_rays
from test_host_traversal import check_smooth_hits, smooth_world

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", ["plain64", "wavefront"])
def test_smooth_mesh_ray_batch(gpu, path):
    world = smooth_world()
    rays = random_rays(40000, [-2.0, 0.05, -1.5], [2.0, 2.0, 1.5], seed=11)
    print("[smooth] ray batch", path, flush=True)
    ho = oracle.trace(world, rays)
    hg = lm.trace_rays(world, rays, path=path)
    r = check_smooth_hits(rays, ho, hg, fp32=(path == "wavefront"))
    print(r)
    assert r["bad_t"] == 0
    assert r["flips"] <= (0 if path == "plain64" else int(5e-4 * r["n"]))


@pytest.fixture(scope="module")
def scene():
    toml, add_meshes = workloads.smooth_mesh_scene(64, 48, 16, 8, 42)
    cfg, cam, world = lm.load_scene(toml)
    add_meshes(world)
    print("[smooth] oracle render", flush=True)
    ref, var, ost = oracle.render(world, cfg)
    return cfg, cam, world, ref, var, ost


def test_smooth_mesh_fp64_matches_oracle(gpu, scene):
    cfg, cam, world, ref, var, ost = scene
    r = lm.render(cfg, cam, world, path="plain64")
    rel = np.linalg.norm(r.rows - ref) / np.linalg.norm(ref)
    print("[smooth] plain64 relRMSE", rel, r.stats, ost, flush=True)
    assert rel < 1e-6
    assert r.stats["extension_rays"] == ost["extension_rays"]
    assert r.stats["shadow_rays"] == ost["shadow_rays"]


def test_smooth_mesh_wavefront_within_3_sigma(gpu, scene):
    cfg, cam, world, ref, var, ost = scene
    r = lm.render(cfg, cam, world, path="wavefront")
    sig = np.sqrt(2.0 * var / 16.0)
    frac = float((np.abs(r.rows - ref) <= 3.0 * sig + 1e-6).mean())
    print("[smooth] wavefront 3-sigma pass", frac, r.stats, flush=True)
    assert frac >= 0.99
    assert abs(r.stats["extension_rays"] - ost["extension_rays"]) <= 0.01 * ost["extension_rays"]


def test_smooth_mesh_wavefront_equals_plain_fp32(gpu, scene):
    cfg, cam, world, _, _, _ = scene
    a = lm.render(cfg, cam, world, path="wavefront")
    b = lm.render(cfg, cam, world, path="plain32")
    np.testing.assert_allclose(a.rows, b.rows, rtol=1e-6, atol=1e-7)
    assert a.stats["extension_rays"] == b.stats["extension_rays"]
