"""The FP32 path's traversal code (4-wide BVH, FP64 islands, transforms,
marched kinds) compiled for the host CPU by tests/native/Makefile and checked
against the oracle on ray batches with the GPU tests' FP32 tolerances: a CPU
check of the device traversal logic (the GPU tests repeat it on the MI355X)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi, workloads
import oracle
from test_gpu_parity import compare_hits, random_rays

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


@pytest.fixture(scope="module")
def harness():
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = C.CDLL(os.path.join(NATIVE, "_build", "libhosttrace.so"))
    L.host_trace_fp32.restype = C.c_int
    L.host_trace_fp32.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(_abi.Ray), C.c_size_t, C.c_uint32,
                                  C.POINTER(_abi.Hit), C.POINTER(C.c_uint32)]
    L.host_trace_error.restype = C.c_char_p
    return L


def host_trace(L, world, rays, exact=True):
    n = len(rays)
    r = lm.rays_to_struct(rays)
    hits = (_abi.Hit * max(n, 1))()
    sw = C.c_uint32()
    rc = L.host_trace_fp32(world.desc, r, n, 1 if exact else 0, hits, C.byref(sw))
    assert rc == 0, L.host_trace_error().decode()
    return lm.hits_to_numpy(hits, n), sw.value


SCENES = {
    "cornell": (lambda: workloads.cornell_box(32, 32, 4, 8), None, [5, 5, -600], [550, 550, 550]),
    "spheres": (lambda: workloads.random_spheres(32, 18, 4, 50, grid=11), None, [-12, 0.01, -12], [12, 3, 12]),
    "mesh": (lambda: workloads.batch_scene(3, seed=5)[0], 12, [-2.5, 0.05, -2.5], [2.5, 3.0, 3.0]),
    "quadrics": (lambda: workloads.quadrics_scene(), None, [-4, 0.05, -3], [4, 3, 3]),
    "traced": (lambda: workloads.traced_scene(), None, [-3.5, 0.05, -2.5], [3.5, 2.0, 2.5]),
}


@pytest.mark.parametrize("scene", list(SCENES))
def test_fp32_traversal_matches_oracle(harness, scene):
    make, freq, lo, hi = SCENES[scene]
    cfg, cam, world = lm.load_scene(make())
    if freq:
        v, f = workloads.displaced_sphere(freq, [0.0, 1.2, 0.0], 1.0, 77)
        world.add_triangles(v, f, material=world.add_material(workloads._microfacet([0.7, 0.6, 0.5], 0.3)))
    rays = random_rays(20000, lo, hi, seed=7)
    ho = oracle.trace(world, rays)
    hh, words = host_trace(harness, world, rays)
    r = compare_hits(rays, ho, hh, fp32=True)
    print(scene, r, "stack words", words)
    assert r["hits"] > 2000
    assert r["bad_t"] == 0
    assert r["flips"] <= int(5e-4 * r["n"])
    assert 0 < words <= 192


@pytest.mark.parametrize("scene", ["cornell", "spheres", "mesh"])
def test_quantised_nodes_contain_their_boxes(harness, scene):
    """The 64-byte nodes' 8-bit child planes (fma(q, step, org) on a per-node
    power-of-two grid) against the FP32 boxes they were quantised from: every
    quantised box contains its FP32 box, and the kernel's child test
    (node4_children, slab distances in fma form from the clamped inverse
    direction) accepts every child the FP32 slab test accepts, for rays from
    inside the scene, from a camera-like distance, axis-parallel rays and
    finite ranges; the quantised boxes stay tight (few extra acceptances)."""
    make, freq, lo, hi = SCENES[scene]
    cfg, cam, world = lm.load_scene(make())
    if freq:
        v, f = workloads.displaced_sphere(freq, [0.0, 1.2, 0.0], 1.0, 77)
        world.add_triangles(v, f, material=world.add_material(workloads._microfacet([0.7, 0.6, 0.5], 0.3)))
    rays = random_rays(1500, lo, hi, seed=11)
    far = random_rays(300, lo, hi, seed=12)
    span = np.asarray(hi, float) - np.asarray(lo, float)
    far[:, :3] = np.asarray(lo, float) - 1.5 * span  # outside the scene, looking in
    far[:, 3:6] = np.asarray(hi, float) - far[:, :3] + np.random.default_rng(3).normal(size=(300, 3)) * span
    par = random_rays(300, lo, hi, seed=13)
    axis = np.random.default_rng(4).integers(0, 3, size=(300, 2))
    par[np.arange(300), 3 + axis[:, 0]] = 0.0  # one or two direction components exactly zero
    par[np.arange(150), 3 + axis[:150, 1]] = 0.0
    par[np.all(par[:, 3:6] == 0, axis=1), 4] = 1.0
    fin = random_rays(300, lo, hi, seed=14)
    fin[:, 7] = np.random.default_rng(5).uniform(0.01, 2.0, 300) * np.linalg.norm(span)
    allr = np.concatenate([rays, far, par, fin]).astype(np.float32).astype(np.float64)
    r = lm.rays_to_struct(allr)
    out = (C.c_ulonglong * 5)()
    L = harness
    L.host_quant_check.restype = C.c_int
    L.host_quant_check.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(_abi.Ray), C.c_size_t,
                                   C.POINTER(C.c_ulonglong)]
    rc = L.host_quant_check(world.desc, r, len(allr), out)
    assert rc == 0, L.host_trace_error().decode()
    tests, fp32_acc, q_acc, lost, not_contained = list(out)
    print(scene, dict(tests=tests, fp32=fp32_acc, quantised=q_acc, lost=lost, not_contained=not_contained))
    assert tests > 20000 and fp32_acc > 1000
    assert not_contained == 0
    assert lost == 0
    # one node level is loose in the random-sphere scene: the root holds the
    # r = 1000 ground sphere beside clusters of r = 0.2 spheres, so its grid
    # step is 2000 / 253 units; every deeper node is tight
    assert q_acc <= (1.8 if scene == "spheres" else 1.05) * fp32_acc


def test_marched_kinds_exact_mode(harness):
    from test_gpu_mesh_batch import MARCHED
    cfg, cam, world = lm.load_scene(MARCHED)
    rng = np.random.default_rng(3)
    n = 20000
    o = rng.uniform([-3, 0.2, 3], [3, 1.5, 5], size=(n, 3))
    tgt = rng.uniform([-1.8, 0.0, -1.0], [1.8, 0.9, 1.0], size=(n, 3))
    rays = np.concatenate([o, tgt - o, np.full((n, 1), 0.001), np.full((n, 1), np.inf)], axis=1)
    ho = oracle.trace(world, rays, exact=True)
    hh, _ = host_trace(harness, world, rays, exact=True)
    r = compare_hits(rays, ho, hh, fp32=True)
    print(r)
    assert r["hits"] > 5000 and r["bad_t"] <= int(2e-4 * n) and r["flips"] <= int(5e-4 * n)


def test_coincident_primitives(harness):
    # 40 identical triangles: coincident centroids force median splits
    tri = 'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n'
    body = "".join('[[triangle]]\nv0 = [0.0, 0.0, 0.0]\nv1 = [1.0, 0.0, 0.0]\nv2 = [0.0, 1.0, 0.0]\n' + tri
                   for _ in range(40))
    _, _, world = lm.load_scene(body)
    rays = np.array([[0.2, 0.2, -1.0, 0.0, 0.0, 1.0, 0.001, np.inf], [2.0, 2.0, -1.0, 0.0, 0.0, 1.0, 0.001, np.inf]])
    hh, _ = host_trace(harness, world, rays)
    assert hh["prim"][0] >= 0 and abs(hh["t"][0] - 1.0) < 1e-6
    assert hh["prim"][1] == -1


def test_superellipsoid_exponent_range(harness):
    # convex (e <= 2) and star-shaped (e > 2) superellipsoids, exponents over
    # the mesh-heavy workload's range U[0.3, 3]
    body = ""
    k = 0
    for e1 in (0.3, 1.0, 2.5):
        for e2 in (0.4, 2.0, 3.0):
            body += (f"[[superellipsoid]]\ncenter = [{2.5 * (k % 3) - 2.5}, {2.5 * (k // 3) - 2.5}, 0.0]\n"
                     f"scale = [1.0, 0.8, 0.9]\ne1 = {e1}\ne2 = {e2}\n"
                     'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n')
            k += 1
    _, _, world = lm.load_scene(body)
    rays = random_rays(30000, [-4.5, -4.5, -4.0], [4.5, 4.5, 4.0], seed=5)
    ho = oracle.trace(world, rays, exact=True)
    hh, _ = host_trace(harness, world, rays, exact=True)
    r = compare_hits(rays, ho, hh, fp32=True)
    print(r)
    assert r["hits"] > 3000
    # the oracle samples 8192 points per chord; the interval walk finds every
    # crossing, so the two differ only on features thinner than that
    assert r["bad_t"] <= int(5e-4 * r["n"]) and r["flips"] <= int(5e-4 * r["n"])


def test_superellipsoid_grazing_rays(harness):
    # rays leaving a superellipsoid's surface at shallow angles (the shadow
    # and bounce rays of its own shading): the sub-chords near the origin lie
    # within 1e-6..1e-3 of the surface, where only the mean-value bound can
    # rule intervals out; outward and inward tilts, convex and star-shaped
    body = ""
    shapes = [(0.36167, 2.18859), (0.3, 0.3), (2.5, 0.4), (1.0, 1.0)]
    for k, (e1, e2) in enumerate(shapes):
        body += (f"[[superellipsoid]]\ncenter = [{3.0 * k}, 0.35, 0.0]\nscale = [0.35, 0.35, 0.35]\n"
                 f"e1 = {e1}\ne2 = {e2}\n" 'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n')
    _, _, world = lm.load_scene(body)
    rng = np.random.default_rng(11)
    n = 4000
    k = rng.integers(0, len(shapes), n)
    c = np.stack([3.0 * k, np.full(n, 0.35), np.zeros(n)], axis=1)
    o = c + rng.normal(size=(n, 3)) * 2.0
    aim = np.concatenate([o, c + rng.uniform(-0.1, 0.1, (n, 3)) - o, np.full((n, 1), 0.001),
                          np.full((n, 1), np.inf)], axis=1)
    h0 = oracle.trace(world, aim, exact=True)
    ok = h0["prim"] >= 0
    p = aim[ok, :3] + aim[ok, 3:6] * h0["t"][ok, None]
    nrm = h0["normal"][ok] * np.where(h0["front_face"][ok], 1.0, -1.0)[:, None]  # outward
    tang = np.cross(nrm, rng.normal(size=nrm.shape))
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    theta = 10.0 ** rng.uniform(-4, -1.3, len(p)) * rng.choice([-1.0, 1.0], len(p))
    d = np.cos(theta)[:, None] * tang + np.sin(theta)[:, None] * nrm
    rays = np.concatenate([p, d, np.full((len(p), 1), 0.001), np.full((len(p), 1), np.inf)], axis=1)
    rays = rays.astype(np.float32).astype(np.float64)
    rays[:, 7] = np.inf
    ho = oracle.trace(world, rays, exact=True)
    hh, _ = host_trace(harness, world, rays, exact=True)
    r = compare_hits(rays, ho, hh, fp32=True)
    print(len(rays), r)
    assert len(rays) > 2500 and r["hits"] > 500
    # a grazing chord may touch the surface between the oracle's 8192 samples
    assert r["bad_t"] <= int(2e-3 * r["n"]) and r["flips"] <= int(5e-3 * r["n"])


def test_pow_pos_accuracy(harness):
    # x^y for x >= 0, y > 0 (the exact-mode superellipsoid's power) against
    # the C library: a few ulp times max(1, |y ln x|)
    L = harness
    L.host_pow_pos.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_size_t]
    rng = np.random.default_rng(4)
    n = 400000
    x = np.concatenate([rng.uniform(0.0, 2.2, n // 2), 2.2 * 10.0 ** rng.uniform(-300, 0, n // 2)])
    y = rng.uniform(0.1, 10.0, n)
    x = np.concatenate([x, [0.0, 1.0, 0.5, 2.0, np.inf, np.nan, 1e-310, 5e-324]])
    y = np.concatenate([y, [2.5, 7.0, 2.0, 3.0, 2.0, 2.0, 0.5, 0.3]])
    out = np.empty_like(x)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    L.host_pow_pos(ptr(x), ptr(y), ptr(out), len(x))
    ref = np.power(x, y)
    assert out[-8] == 0.0 and out[-7] == 1.0
    assert out[-6] == pytest.approx(0.25, rel=4 * 2.0 ** -52) and out[-5] == pytest.approx(8.0, rel=4 * 2.0 ** -52)
    assert np.isinf(out[-4]) and np.isnan(out[-3])
    fin = np.isfinite(ref) & (ref > 1e-300)
    lnx = np.log(np.where(x > 0, x, 1.0))
    err = np.abs(out[fin] - ref[fin]) / ref[fin] / (np.maximum(1.0, np.abs(y * lnx)[fin]) * 2.0 ** -52)
    print("max error (ulp * max(1, |y ln x|))", err.max())
    assert err.max() < 4.0
    tiny = ~fin & np.isfinite(ref)
    assert np.all(np.abs(out[tiny] - ref[tiny]) <= 1e-300)


def test_parallel_bvh_build_gives_the_same_hits(harness, monkeypatch):
    """The BVH of a large mesh is built on worker threads below its top
    levels; the tree, and so every hit, is the one-thread build's."""
    cfg, cam, world = lm.load_scene(workloads.batch_scene(3, seed=5)[0])
    v, f = workloads.displaced_sphere(64, [0.0, 1.2, 0.0], 1.0, 77)   # 81,920 triangles
    world.add_triangles(v, f, material=world.add_material(workloads._microfacet([0.7, 0.6, 0.5], 0.3)))
    rays = random_rays(20000, [-2.5, 0.05, -2.5], [2.5, 3.0, 3.0], seed=11)
    monkeypatch.setenv("LMN_BUILD_THREADS", "1")
    one, _ = host_trace(harness, world, rays)
    monkeypatch.setenv("LMN_BUILD_THREADS", "8")
    many, _ = host_trace(harness, world, rays)
    assert (one["prim"] >= 0).sum() > 2000
    assert np.array_equal(one["prim"], many["prim"]) and np.array_equal(one["t"], many["t"])
    assert np.array_equal(one["normal"], many["normal"])


def smooth_world():
    toml, add_meshes = workloads.smooth_mesh_scene()
    cfg, cam, world = lm.load_scene(toml)
    assert add_meshes(world) == 4000
    return world


def check_smooth_hits(rays, ho, hx, fp32):
    # interpolated normals and texture coordinates (MeshTriangle::hit,
    # src/obj.rs:260-313) on the hits where both sides found the same
    # triangle; the mesh triangles are prims 2..4001
    r = compare_hits(rays, ho, hx, fp32=fp32)
    same = (ho["prim"] == hx["prim"]) & (ho["prim"] >= 2)
    assert same.sum() > 2000
    dots = np.einsum("ij,ij->i", ho["normal"][same], hx["normal"][same])
    du = np.abs(ho["u"][same] - hx["u"][same])
    dv = np.abs(ho["v"][same] - hx["v"][same])
    if fp32:
        assert (dots > 1 - 1e-5).mean() > 0.999 and dots.min() > 0.99
        assert (np.maximum(du, dv) < 1e-4).mean() > 0.999
    else:
        assert dots.min() > 1 - 1e-12
        assert np.maximum(du, dv).max() < 1e-9
    assert (ho["front_face"][same] == hx["front_face"][same]).mean() > 0.999
    return r


def test_smooth_mesh_attributes(harness):
    world = smooth_world()
    rays = random_rays(20000, [-2.0, 0.05, -1.5], [2.0, 2.0, 1.5], seed=11)
    ho = oracle.trace(world, rays)
    hh, _ = host_trace(harness, world, rays)
    r = check_smooth_hits(rays, ho, hh, fp32=True)
    print(r)
    assert r["bad_t"] == 0
    assert r["flips"] <= int(5e-4 * r["n"])
