"""Disk, cylinder, cone, ellipsoid, hemisphere, annulus and capsule
(src/disk.rs, src/cylinder.rs, src/cone.rs, src/ellipsoid.rs,
src/hemisphere.rs, src/annulus.rs, src/capsule.rs) in the oracle: the
reference's own tests restated, and closed forms for the hit distance,
normal and (u, v) they fix."""
import math

import numpy as np
import pytest

import luminara_amd as lm
import oracle

MAT = 'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }'


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def world_of(body):
    toml = f"""
[render]
width = 8
height = 8
samples = 1
max_depth = 2

[camera]
look_from = [0.0, 0.0, 10.0]
look_at = [0.0, 0.0, 0.0]
vfov = 40.0

{body}
"""
    return lm.load_scene(toml)[2]


def trace(world, o, d):
    rays = np.array([[*o, *d, 0.001, np.inf]], dtype=np.float64)
    return oracle.trace(world, rays)[0]


def test_disk_reference_tests_and_closed_forms():
    w = world_of(f"[[disk]]\ncenter = [0.0, 0.0, 0.0]\nnormal = [0.0, 1.0, 0.0]\nradius = 1.0\n{MAT}\n")
    h = trace(w, [0, 1, 0], [0, -1, 0])            # test_disk_hit
    assert h["prim"] == 0 and h["t"] == 1.0
    assert list(h["normal"]) == [0.0, 1.0, 0.0] and h["front_face"] == 1
    assert trace(w, [5, 1, 0], [0, -1, 0])["prim"] == -1   # test_disk_miss_outside_radius
    # (u, v) = (atan2(z, x) / 2pi + 0.5, r / R): the point (0.5, 0, 0.5)
    h = trace(w, [0.5, 2, 0.5], [0, -1, 0])
    assert h["u"] == pytest.approx(math.atan2(0.5, 0.5) / (2 * math.pi) + 0.5, rel=1e-15)
    assert h["v"] == pytest.approx(math.sqrt(0.5), rel=1e-15)
    # from below: the normal faces the ray; a zero normal becomes (0, 1, 0)
    h = trace(w, [0.2, -3, 0.1], [0, 1, 0])
    assert h["t"] == 3.0 and list(h["normal"]) == [0.0, -1.0, 0.0] and h["front_face"] == 0
    w0 = world_of(f"[[disk]]\ncenter = [0.0, 0.0, 0.0]\nnormal = [0.0, 0.0, 0.0]\nradius = 1.0\n{MAT}\n")
    assert trace(w0, [0, 1, 0], [0, -1, 0])["t"] == 1.0


def test_cylinder_reference_tests_and_closed_forms():
    w = world_of(f"[[cylinder]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\nheight = 2.0\n{MAT}\n")
    h = trace(w, [3, 1, 0], [-1, 0, 0])             # test_cylinder_hit: point.x = 1
    assert h["prim"] == 0 and 3.0 - h["t"] == pytest.approx(1.0, abs=1e-12)
    assert np.allclose(h["normal"], [1, 0, 0]) and h["front_face"] == 1
    assert h["v"] == pytest.approx(0.5) and h["u"] == pytest.approx(0.5)   # theta = atan2(-0, 1) + pi
    assert trace(w, [3, 5, 0], [-1, 0, 0])["prim"] == -1   # test_cylinder_miss_above
    # open tube: from inside the far wall, facing inward
    h = trace(w, [0, 1, 0], [1, 0, 0])
    assert h["t"] == pytest.approx(1.0) and np.allclose(h["normal"], [-1, 0, 0]) and h["front_face"] == 0
    # a ray through the top misses the open end, then meets the wall
    h = trace(w, [0.0, 3.0, 0.0], [0.5, -1.0, 0.0])
    assert h["t"] == pytest.approx(2.0) and h["v"] == pytest.approx(0.5)


def test_cone_reference_tests_and_closed_forms():
    w = world_of(f"[[cone]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\nheight = 2.0\n{MAT}\n")
    h = trace(w, [3, 0.5, 0], [-1, 0, 0])           # test_cone_hit: r(y) = (2 - y) / 2
    assert h["prim"] == 0 and h["t"] == pytest.approx(3.0 - 0.75, abs=1e-12)
    n = np.array([1.0, 0.5, 0.0]) / math.sqrt(1.25)   # (x / r(y), r / h, z / r(y)), normalised
    assert np.allclose(h["normal"], n) and h["v"] == pytest.approx(0.25)
    assert trace(w, [3, 5, 0], [-1, 0, 0])["prim"] == -1   # test_cone_miss_above_apex
    # the upper nappe above the apex is outside [y_min, y_max]
    assert trace(w, [3, 3, 0], [-1, 0, 0])["prim"] == -1


def test_emissive_disk_is_an_nee_light():
    w = world_of('[[disk]]\ncenter = [0.0, 2.0, 0.0]\nnormal = [0.0, -2.0, 0.0]\nradius = 0.5\n'
                 'material = { type = "emissive", color = [1.0, 0.5, 0.25], intensity = 4.0 }\n')
    d = w.desc_struct
    assert d.n_lights == 1 and d.lights[0].kind == 2
    assert list(d.lights[0].normal) == [0.0, -1.0, 0.0] and list(d.lights[0].emission) == [4.0, 2.0, 1.0]
    assert d.lights[0].radius == 0.5


def sphere_uv(n):
    return (math.atan2(-n[2], n[0]) + math.pi) / (2 * math.pi), math.acos(-n[1]) / math.pi


def test_ellipsoid_reference_tests_and_closed_forms():
    w = world_of(f"[[ellipsoid]]\ncenter = [0.0, 0.0, 0.0]\nradii = [2.0, 1.0, 1.0]\n{MAT}\n")
    h = trace(w, [0, 0, -5], [0, 0, 1])              # test_ellipsoid_hit
    assert h["prim"] == 0 and h["t"] == 4.0 and list(h["normal"]) == [0.0, 0.0, -1.0]
    # a surface point P = (2 cos a, sin a cos b, sin a sin b) reached along an
    # arbitrary direction: t = |P - o|, normal = unit(P.x / 4, P.y, P.z)
    a, b = 0.7, 0.4
    P = np.array([2 * math.cos(a), math.sin(a) * math.cos(b), math.sin(a) * math.sin(b)])
    n = np.array([P[0] / 4, P[1], P[2]])
    n /= np.linalg.norm(n)
    o = P + 3.0 * (n + np.array([0.2, -0.1, 0.3]))
    d = (P - o) / 3.0
    h = trace(w, o, d)
    assert h["t"] == pytest.approx(3.0, rel=1e-12) and np.allclose(h["normal"], n, atol=1e-12)
    assert (h["u"], h["v"]) == pytest.approx(sphere_uv(n), rel=1e-12)
    # test_ellipsoid_bounding_box: a ray just outside centre +/- radii misses
    w = world_of(f"[[ellipsoid]]\ncenter = [1.0, 2.0, 3.0]\nradii = [0.5, 1.0, 0.3]\n{MAT}\n")
    assert trace(w, [0.49, 2.0, -5.0], [0, 0, 1])["prim"] == -1
    assert trace(w, [0.51, 2.0, -5.0], [0, 0, 1])["prim"] == 0
    assert trace(w, [1.0, 3.01, -5.0], [0, 0, 1])["prim"] == -1


def test_hemisphere_reference_tests_and_closed_forms():
    w = world_of(f"[[hemisphere]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\n{MAT}\n")
    h = trace(w, [0, 5, 0], [0, -1, 0])              # hemisphere_hit_from_above
    assert h["prim"] == 0 and h["t"] == 4.0 and h["front_face"] == 1
    # hemisphere_miss_below: the near root (y = -1) is cut away, the far one
    # (y = +1) is hit from inside the dome
    h = trace(w, [0, -5, 0], [0, 1, 0])
    assert h["t"] == 6.0 and list(h["normal"]) == [0.0, -1.0, 0.0] and h["front_face"] == 0
    # a horizontal ray below the rim misses; just above it hits
    assert trace(w, [-5, -0.01, 0], [1, 0, 0])["prim"] == -1
    h = trace(w, [-5, 0.6, 0], [1, 0, 0])
    assert h["t"] == pytest.approx(5 - 0.8, rel=1e-12) and np.allclose(h["normal"], [-0.8, 0.6, 0])
    assert (h["u"], h["v"]) == pytest.approx(sphere_uv([-0.8, 0.6, 0.0]), rel=1e-12)
    # hemisphere_bbox_upper_only: centre (0, 2, 0), r = 1 spans y in [2, 3]
    w = world_of(f"[[hemisphere]]\ncenter = [0.0, 2.0, 0.0]\nradius = 1.0\n{MAT}\n")
    assert trace(w, [0, 2.5, -5], [0, 0, 1])["prim"] == 0
    assert trace(w, [0, 1.9, -5], [0, 0, 1])["prim"] == -1


def test_annulus_reference_tests_and_closed_forms():
    body = ("[[annulus]]\ncenter = [0.0, 0.0, 0.0]\nnormal = [0.0, 1.0, 0.0]\ninner_radius = 0.5\n"
            f"outer_radius = 1.0\n{MAT}\n")
    w = world_of(body)
    h = trace(w, [0.75, 5, 0], [0, -1, 0])           # annulus_hit_in_ring
    assert h["prim"] == 0 and h["t"] == 5.0 and list(h["normal"]) == [0.0, 1.0, 0.0]
    assert h["u"] == pytest.approx(0.5) and h["v"] == pytest.approx(0.5)   # atan2(0, .75) = 0; (0.75 - .5) / .5
    assert trace(w, [0, 5, 0], [0, -1, 0])["prim"] == -1      # annulus_miss_in_hole
    assert trace(w, [2, 5, 0], [0, -1, 0])["prim"] == -1      # annulus_miss_outside
    # tilted: normal (0, 3, 4) / 5 given unnormalised; a point at radius 0.8 on the ring
    w = world_of(body.replace("normal = [0.0, 1.0, 0.0]", "normal = [0.0, 3.0, 4.0]"))
    e1 = np.array([1.0, 0.0, 0.0])
    e2 = np.array([0.0, 4.0, -3.0]) / 5.0
    P = 0.8 * (math.cos(1.1) * e1 + math.sin(1.1) * e2)
    n = np.array([0.0, 0.6, 0.8])
    h = trace(w, P + 2.0 * n, -n)
    assert h["t"] == pytest.approx(2.0, rel=1e-12) and np.allclose(h["normal"], n)
    assert h["v"] == pytest.approx(0.6, rel=1e-12)
    assert h["u"] == pytest.approx(math.atan2(P[2], P[0]) / (2 * math.pi) + 0.5, rel=1e-12)


def test_capsule_reference_tests_and_closed_forms():
    w = world_of(f"[[capsule]]\ncenter = [0.0, 0.0, 0.0]\nradius = 0.5\nheight = 2.0\n{MAT}\n")
    h = trace(w, [0, 2, -3], [0, 0, 1])              # test_capsule_hit: the top sphere
    assert h["prim"] == 2 and h["t"] == 2.5
    assert trace(w, [5, 5, -3], [0, 0, 1])["prim"] == -1     # miss entirely
    h = trace(w, [0, 1, -3], [0, 0, 1])                         # the body
    assert h["prim"] == 0 and h["t"] == 2.5 and np.allclose(h["normal"], [0, 0, -1])
    h = trace(w, [0, -0.3, -3], [0, 0, 1])                      # the bottom sphere below the body
    assert h["prim"] == 1 and h["t"] == pytest.approx(3 - 0.4, rel=1e-12)
    # straight down the axis: the open body is passed through, the top sphere's cap is hit
    h = trace(w, [0, 5, 0], [0, -1, 0])
    assert h["prim"] == 2 and h["t"] == 2.5
