"""Disk, cylinder, cone, ellipsoid, hemisphere, annulus, capsule, prism,
wedge and pyramid (src/disk.rs, src/cylinder.rs, src/cone.rs,
src/ellipsoid.rs, src/hemisphere.rs, src/annulus.rs, src/capsule.rs,
src/prism.rs, src/wedge.rs, src/pyramid.rs) and the sphere-traced rounded
box and Mobius strip (src/rounded_box.rs, src/mobius.rs) in the oracle: the
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


def trace(world, o, d, exact=True):
    rays = np.array([[*o, *d, 0.001, np.inf]], dtype=np.float64)
    return oracle.trace(world, rays, exact=exact)[0]


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


def test_prism_reference_tests_and_closed_forms():
    w = world_of(f"[[prism]]\ncenter = [0.0, 0.0, 0.0]\nside = 2.0\nheight = 3.0\n{MAT}\n")
    r = 2.0 / math.sqrt(3.0)                          # circumradius; front vertex at z = r
    h = trace(w, [0, 1.5, 5], [0, 0, -1])             # test_prism_hit_side (the front edge)
    assert h["prim"] in (2, 3, 6, 7) and h["t"] == pytest.approx(5 - r, rel=1e-12)
    assert trace(w, [0, 10, 5], [0, 0, -1])["prim"] == -1      # test_prism_miss (side 2 here)
    # test_prism_bounding_box: y in [0, 3]
    assert trace(w, [0, 2.99, 5], [0, 0, -1])["prim"] >= 0
    assert trace(w, [0, 3.01, 5], [0, 0, -1])["prim"] == -1
    # the back side (v1 -> v2, z = -r/2): the reference's normal edge x up = (0, 0, 1)
    # points into the solid, so a ray from behind meets its back (front_face false)
    h = trace(w, [0.3, 1.0, -5], [0, 0, 1])
    assert h["prim"] in (4, 5) and h["t"] == pytest.approx(5 - r / 2, rel=1e-12)
    assert np.allclose(h["normal"], [0, 0, -1]) and h["front_face"] == 0
    # top cap (top0, top2, top1), normal +y; (u, v) barycentric in that order
    h = trace(w, [0.1, 10, 0], [0, -1, 0])
    assert h["prim"] == 1 and h["t"] == pytest.approx(7.0, rel=1e-14)
    assert list(h["normal"]) == [0.0, 1.0, 0.0] and h["front_face"] == 1
    a, b, c = np.array([0, 3, r]), np.array([1, 3, -r / 2]), np.array([-1, 3, -r / 2])
    uv = np.linalg.lstsq(np.stack([b - a, c - a], 1), np.array([0.1, 3, 0]) - a, rcond=None)[0]
    assert (h["u"], h["v"]) == pytest.approx(tuple(uv), abs=1e-12)


def test_wedge_reference_tests_and_closed_forms():
    w = world_of(f"[[wedge]]\nmin = [-1.0, 0.0, 0.0]\nmax = [1.0, 2.0, 4.0]\n{MAT}\n")
    h = trace(w, [0, 5, 2], [0, -1, 0])               # test_wedge_hit_ramp: y = 1 at z = 2
    ramp_n = np.array([0, 1, -0.5]) / math.sqrt(1.25)
    assert h["prim"] == 4 and h["t"] == pytest.approx(4.0, rel=1e-12)
    assert np.allclose(h["normal"], ramp_n) and h["front_face"] == 1
    assert (h["u"], h["v"]) == pytest.approx((0.5, 0.5))
    h = trace(w, [0, -1, 2], [0, 1, 0])               # test_wedge_hit_bottom
    assert h["prim"] == 0 and h["t"] == 1.0 and list(h["normal"]) == [0.0, -1.0, 0.0]
    assert trace(w, [5, 5, 5], [1, 1, 1])["prim"] == -1         # test_wedge_miss
    # test_wedge_bounding_box and the faces: back (z = 4), left under the ramp
    h = trace(w, [0, 1, 9], [0, 0, -1])
    assert h["prim"] == 1 and h["t"] == 5.0 and list(h["normal"]) == [0.0, 0.0, 1.0]
    assert (h["u"], h["v"]) == pytest.approx((0.5, 0.5))
    h = trace(w, [-5, 0.5, 3], [1, 0, 0])
    assert h["prim"] == 2 and h["t"] == 4.0 and list(h["normal"]) == [-1.0, 0.0, 0.0]
    assert (h["u"], h["v"]) == pytest.approx((0.75, 0.25))
    assert trace(w, [-5, 1.9, 3], [1, 0, 0])["prim"] == -1       # above the ramp at z = 3
    assert trace(w, [0, 2.01, 9], [0, 0, -1])["prim"] == -1


def test_pyramid_reference_tests_and_closed_forms():
    w = world_of(f"[[pyramid]]\ncenter = [0.0, 0.0, 0.0]\nbase_size = 2.0\nheight = 3.0\n{MAT}\n")
    h = trace(w, [0, 1, 5], [0, 0, -1])               # test_pyramid_hit_side: z = 1 - y / 3
    assert h["prim"] == 4 and h["t"] == pytest.approx(5 - 2 / 3, rel=1e-12)
    # its normal unit(e1 x e2) of (b2, b3, apex) is (0, -1, -3) / sqrt(10): inward
    assert np.allclose(h["normal"], np.array([0, 1, 3]) / math.sqrt(10)) and h["front_face"] == 0
    h = trace(w, [0.3, -1, -0.2], [0, 1, 0])          # test_pyramid_hit_base, off the diagonal
    assert h["prim"] == 0 and h["t"] == 1.0 and list(h["normal"]) == [0.0, -1.0, 0.0] and h["front_face"] == 0
    assert trace(w, [5, 5, 5], [1, 1, 1])["prim"] == -1          # test_pyramid_miss


RBOX = "[[rounded_box]]\ncenter = [0.0, 0.0, 0.0]\nhalf_size = [1.0, 1.0, 1.0]\nradius = 0.2\n" + MAT + "\n"


@pytest.mark.parametrize("exact", [True, False])
def test_rounded_box_reference_tests(exact):
    w = world_of(RBOX)
    h = trace(w, [0, 0, -5], [0, 0, 1], exact)        # test_rounded_box_hit: t ~ 5 - 1.2
    assert h["prim"] == 0 and abs(h["t"] - 3.8) < 0.1
    assert np.allclose(h["normal"], [0, 0, -1], atol=1e-4) and h["front_face"] == 1
    small = "[[rounded_box]]\ncenter = [0.0, 0.0, 0.0]\nhalf_size = [0.5, 0.5, 0.5]\nradius = 0.1\n" + MAT
    assert trace(world_of(small), [5, 5, -3], [0, 0, 1], exact)["prim"] == -1   # test_rounded_box_miss
    # test_rounded_box_bbox: centre (1, 2, 3), extent 0.6 -> x in [0.4, 1.6]
    bb = small.replace("center = [0.0, 0.0, 0.0]", "center = [1.0, 2.0, 3.0]")
    assert trace(world_of(bb), [1.59, 2.59, -5], [0, 0, 1], exact)["prim"] == -1  # past the rounded corner
    assert trace(world_of(bb), [1.59, 2.0, -5], [0, 0, 1], exact)["prim"] == 0    # the rounded edge
    assert trace(world_of(bb), [1.61, 2.0, -5], [0, 0, 1], exact)["prim"] == -1


def test_rounded_box_exact_closed_forms():
    w = world_of(RBOX)
    # face: the plane z = -1.2; edge: the cylinder of radius 0.2 about x = 1, z = -1
    # (a ray along z at x = 1.1 meets it at z = -1 - sqrt(0.04 - 0.01)); corner: the
    # sphere of radius 0.2 about (1, 1, -1) along the diagonal
    h = trace(w, [0.3, -0.4, -5], [0, 0, 2])
    assert h["t"] == pytest.approx(3.8 / 2, rel=1e-12) and np.allclose(h["normal"], [0, 0, -1], atol=1e-6)
    assert (h["u"], h["v"]) == pytest.approx(((0.3 / 1.2 + 1) / 2, (-0.4 / 1.2 + 1) / 2), abs=1e-9)
    h = trace(w, [1.1, 0.0, -5], [0, 0, 1])
    assert h["t"] == pytest.approx(4 - math.sqrt(0.03), rel=1e-12)
    assert np.allclose(h["normal"], [0.1 / 0.2, 0, -math.sqrt(0.03) / 0.2], atol=1e-5)
    d = np.array([-1.0, -1.0, 1.0]) / math.sqrt(3)
    o = np.array([1.0, 1.0, -1.0]) - 4.0 * d
    h = trace(w, o, d)
    assert h["t"] == pytest.approx(4.0 - 0.2, rel=1e-12) and np.allclose(h["normal"], -d, atol=1e-5)
    # from inside: exact mode finds the exit, the reference's trace stops at t_min
    h = trace(w, [0, 0, 0], [1, 0, 0])
    assert h["t"] == pytest.approx(1.2, rel=1e-12) and h["front_face"] == 0
    assert trace(w, [0, 0, 0], [1, 0, 0], exact=False)["t"] == 0.001


def test_rounded_box_reference_trace_steps():
    # the reference's trace steps max(dist, 1e-4) in t with the direction
    # unnormalised, so a step of dist moves |d| dist: with |d| = 1.5 the first
    # step from t_min lands inside the box (dist < 0 there is taken as the
    # hit); with |d| = 4 it jumps over the box, a miss.  Exact mode hits the
    # face at z = -1.2 either way.
    w = world_of(RBOX)
    h = trace(w, [0, 0, -5], [0, 0, 1.5], exact=False)
    assert h["prim"] == 0 and h["t"] == pytest.approx(0.001 + (5 - 0.0015 - 1.2), rel=1e-12)
    assert trace(w, [0, 0, -5], [0, 0, 4], exact=False)["prim"] == -1
    assert trace(w, [0, 0, -5], [0, 0, 4])["t"] == pytest.approx(3.8 / 4, rel=1e-12)
    assert trace(w, [0, 0, -5], [0, 0, 1.5])["t"] == pytest.approx(3.8 / 1.5, rel=1e-12)


MOB = "[[mobius]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\nwidth = 0.3\nthickness = 0.05\n" + MAT + "\n"


def test_mobius_reference_tests_and_geometry():
    w = world_of(MOB)
    h = trace(w, [1, 0, -3], [0, 0, 1])               # test_mobius_hit
    assert h["prim"] == 0
    assert trace(w, [0, 5, -3], [0, 0, 1])["prim"] == -1     # test_mobius_miss
    # a ray down the y axis at the centre passes through the hole
    assert trace(w, [0, 3, 0], [0, -1, 0])["prim"] == -1
    # the half twist: the ribbon's direction at angle u is (cos u cos u/2,
    # sin u/2, sin u cos u/2): radial and flat at angle 0 (the segment
    # x in [0.7, 1.3] at y = z = 0), vertical at angle pi (y in [-0.3, 0.3] at
    # x = -1, z = 0).  Thickened by 0.05, the flat ribbon is met from above at
    # y = 0.05, the vertical one across at x = -1.05; a ray along x at
    # y = 0.2 passes 0.15 over the flat ribbon and meets the vertical one on
    # its inner side, x = -0.95 (rays in the plane z = 0, where the distance
    # estimate is exact: accepted within 1e-4)
    h = trace(w, [1.0, 3.0, 0.0], [0, -1, 0])
    assert h["prim"] == 0 and abs((3 - h["t"]) - 0.05) <= 1e-4
    h = trace(w, [-3.0, 0.2, 0.0], [1, 0, 0])
    assert h["prim"] == 0 and abs((-3 + h["t"]) + 1.05) <= 1e-4
    assert h["normal"][0] < -0.99   # (the z difference straddles the branch cut of atan2 at angle pi)
    h = trace(w, [3.0, 0.2, 0.0], [-1, 0, 0])
    assert h["prim"] == 0 and abs((3 - h["t"]) + 0.95) <= 1e-4
    _, _, wd = lm.load_scene(MOB.replace("thickness = 0.05\n", ""))
    assert list(wd.prims[0].p[:6]) == [0.0, 0.0, 0.0, 1.0, 0.3, 0.05]
