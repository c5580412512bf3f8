"""Scene loader (host C++, no GPU): the reference's load_scene tests restated,
its defaults, light rules and ordering, PLY meshes, and a parse sweep over
the reference's example scenes when they are present."""
import glob
import os

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi, workloads

REF_SCENES = "/root/reference/examples/scenes"


def test_load_simple_scene():  # src/scene.rs:2222-2256
    toml = """
[render]
width = 400
height = 200
samples = 16
max_depth = 10

[camera]
look_from = [0.0, 1.0, 5.0]
look_at = [0.0, 0.0, 0.0]
vfov = 60.0

[[sphere]]
center = [0.0, 0.0, 0.0]
radius = 1.0
[sphere.material]
type = "lambertian"
color = [0.8, 0.2, 0.1]

[[sphere]]
center = [2.0, 0.0, 0.0]
radius = 0.5
[sphere.material]
type = "metal"
color = [0.9, 0.9, 0.9]
fuzz = 0.1
"""
    cfg, cam, world = lm.load_scene(toml)
    assert (cfg.width, cfg.height, cfg.samples_per_pixel, cfg.max_depth) == (400, 200, 16, 10)
    mats = world.materials()
    assert [m.kind for m in mats] == [_abi.MAT_LAMBERTIAN, _abi.MAT_METAL]
    assert abs(mats[1].fuzz - 0.1) < 1e-15
    # the dotted [sphere.material] header addresses the last [[sphere]]
    assert [p.material for p in world.prims] == [0, 1]


def test_materials_created_from_toml():  # src/scene.rs:2258-2277
    toml = """
[[sphere]]
center = [0.0, 0.0, 0.0]
radius = 1.0
[sphere.material]
type = "dielectric"
refraction_index = 1.5

[[sphere]]
center = [3.0, 0.0, 0.0]
radius = 1.0
[sphere.material]
type = "emissive"
color = [1.0, 1.0, 1.0]
intensity = 5.0
"""
    _, _, world = lm.load_scene(toml)
    assert len(world.prims) == 2
    # an emissive sphere is registered as a sphere light (src/scene.rs:1182-1202)
    lights = world.lights
    assert len(lights) == 1 and lights[0].kind == _abi.LIGHT_SPHERE
    assert list(lights[0].emission) == [5.0, 5.0, 5.0]


def test_invalid_toml_returns_error():  # src/scene.rs:2279-2288
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene("this is {{{{ not valid TOML at all")
    assert e.value.code == _abi.LMN_ERR_PARSE
    assert "TOML parse error" in str(e.value)


def test_missing_material_type_returns_error():  # src/scene.rs:2290-2302
    with pytest.raises(_abi.LuminaraError):
        lm.load_scene("[[sphere]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\n[sphere.material]\ncolor = [1.0, 0.0, 0.0]\n")


def test_geometry_outside_the_path_is_rejected_by_name():  # src/scene.rs:2304-2392 (scope subset)
    base = '[[sphere]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\nmaterial = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n'
    _, _, world = lm.load_scene(base)
    assert len(world.prims) == 1
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene(base + '[[fog]]\ncenter = [0.0, 1.0, 0.0]\nradius = 2.0\ndensity = 0.5\n'
                             'color = [0.8, 0.8, 0.8]\n')
    assert e.value.code == _abi.LMN_ERR_UNSUPPORTED and "fog" in str(e.value)
    # cylinders, cones and disks are on the path, after spheres in load order
    _, _, world = lm.load_scene(base + '[[cylinder]]\ncenter = [0.0, 0.0, 0.0]\nradius = 0.5\nheight = 1.0\n'
                                       'material = { type = "metal", color = [0.5, 0.5, 0.5] }\n')
    assert [p.kind for p in world.prims] == [_abi.PRIM_SPHERE, _abi.PRIM_CYLINDER]


def test_capsule_ellipsoid_hemisphere_annulus_load_order():  # src/scene.rs:1308-1386, :2341-2385
    mat = 'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n'
    body = ('[[annulus]]\ncenter = [0.0, 0.0, 0.0]\nnormal = [0.0, 1.0, 0.0]\ninner_radius = 0.5\n'
            'outer_radius = 1.0\n' + mat +
            '[[hemisphere]]\ncenter = [0.0, 0.0, 0.0]\nradius = 1.0\n' + mat +
            '[[ellipsoid]]\ncenter = [0.0, 0.0, 0.0]\nradii = [1.0, 0.5, 0.7]\n' + mat +
            '[[capsule]]\ncenter = [1.0, 2.0, 3.0]\nradius = 0.3\nheight = 1.0\n' + mat)
    _, _, world = lm.load_scene(body)
    # fixed type order: capsule, ellipsoid, (light), hemisphere, ..., annulus;
    # a capsule is one world-list entry of a cylinder and two spheres
    assert [p.kind for p in world.prims] == [_abi.PRIM_CYLINDER, _abi.PRIM_SPHERE, _abi.PRIM_SPHERE,
                                             _abi.PRIM_ELLIPSOID, _abi.PRIM_HEMISPHERE, _abi.PRIM_ANNULUS]
    assert [p.object for p in world.prims] == [0, 0, 0, 1, 2, 3]
    cyl, bottom, top = world.prims[:3]
    assert list(cyl.p[:6]) == [1.0, 2.0, 3.0, 0.3, 2.0, 3.0]
    assert list(bottom.p[:4]) == [1.0, 2.0, 3.0, 0.3] and list(top.p[:4]) == [1.0, 3.0, 3.0, 0.3]
    assert list(world.prims[5].p[:8]) == [0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.5, 1.0]
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene('[[ellipsoid]]\ncenter = [0.0, 0.0, 0.0]\nradii = [1.0, 0.5, 0.7]\nbump_strength = 0.2\n' + mat)
    assert e.value.code == _abi.LMN_ERR_UNSUPPORTED and "bump" in str(e.value)


def test_defaults_match_render_config_default():  # src/render.rs:422-431
    cfg, cam, world = lm.load_scene("")
    assert (cfg.width, cfg.height, cfg.samples_per_pixel, cfg.max_depth, cfg.seed) == (800, 450, 100, 50, 31337)
    assert len(world.prims) == 0


def test_fixed_type_order_and_light_rules():
    # objects enter the world by type, not by file order (src/scene.rs:1180-1642):
    # spheres before planes before quads before boxes
    toml = """
[[box]]
min = [0.0, 0.0, 0.0]
max = [1.0, 1.0, 1.0]
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
[[quad]]
q = [0.0, 3.0, 0.0]
u = [1.0, 0.0, 0.0]
v = [0.0, 0.0, 1.0]
material = { type = "emissive", color = [1.0, 0.9, 0.8], intensity = 4.0 }
[[plane]]
point = [0.0, 0.0, 0.0]
normal = [0.0, 1.0, 0.0]
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
[[sphere]]
center = [0.0, 5.0, 0.0]
radius = 0.5
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
"""
    _, _, world = lm.load_scene(toml)
    kinds = [p.kind for p in world.prims]
    assert kinds[:3] == [_abi.PRIM_SPHERE, _abi.PRIM_PLANE, _abi.PRIM_QUAD]
    assert kinds[3:] == [_abi.PRIM_RECT_XY] * 2 + [_abi.PRIM_RECT_XZ] * 2 + [_abi.PRIM_RECT_YZ] * 2
    lights = world.lights
    assert len(lights) == 1 and lights[0].kind == _abi.LIGHT_RECT
    np.testing.assert_allclose(list(lights[0].emission), [4.0, 3.6, 3.2])


def test_camera_keeps_toml_aspect_under_overrides():
    # -w/-H overrides change the film, not the camera (src/main.rs:205-210 vs
    # src/scene.rs:1146-1174): the camera frame depends on the TOML size only
    toml = "[render]\nwidth = 200\nheight = 100\n[camera]\nlook_from = [0.0, 0.0, 0.0]\nlook_at = [0.0, 0.0, -1.0]\nvfov = 90.0\n"
    cfg, cam, _ = lm.load_scene(toml)
    h = np.array(list(cam.horizontal))
    v = np.array(list(cam.vertical))
    assert abs(np.linalg.norm(h) / np.linalg.norm(v) - 2.0) < 1e-12


def _ply(verts, faces):
    lines = ["ply", "format ascii 1.0", f"element vertex {len(verts)}", "property float x", "property float y",
             "property float z", f"element face {len(faces)}", "property list uchar int vertex_indices",
             "end_header"]
    lines += [" ".join(repr(float(x)) for x in v) for v in verts]
    lines += [" ".join([str(len(f))] + [str(i) for i in f]) for f in faces]
    return "\n".join(lines) + "\n"


def test_ply_fan_triangulation_scale_and_offset(tmp_path):  # src/obj.rs:157-246
    verts = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    path = tmp_path / "q.ply"
    path.write_text(_ply(verts, [[0, 1, 2, 3]]))
    toml = (f'[[mesh]]\nfile = "{path}"\nscale = 2.0\noffset = [1.0, 0.0, 0.0]\n'
            'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n')
    _, _, world = lm.load_scene(toml)
    tris = [np.array(list(p.p)[:9]).reshape(3, 3) for p in world.prims]
    assert len(tris) == 2  # a quad face fans into (0,1,2), (0,2,3)
    np.testing.assert_allclose(tris[0], [[1, 0, 0], [3, 0, 0], [3, 2, 0]])
    np.testing.assert_allclose(tris[1], [[1, 0, 0], [3, 2, 0], [1, 2, 0]])


def test_ply_errors(tmp_path):
    bad = tmp_path / "bad.ply"
    bad.write_text("not a ply\n")
    toml = f'[[mesh]]\nfile = "{bad}"\nmaterial = {{ type = "lambertian", color = [0.5, 0.5, 0.5] }}\n'
    with pytest.raises(_abi.LuminaraError):
        lm.load_scene(toml)
    oob = tmp_path / "oob.ply"
    oob.write_text(_ply([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 5]]))
    with pytest.raises(_abi.LuminaraError):
        lm.load_scene(toml.replace(str(bad), str(oob)))
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene(toml.replace(str(bad), str(tmp_path / "missing.ply")))
    assert e.value.code == _abi.LMN_ERR_IO


def test_ply_and_binary_meshes_agree(tmp_path):
    toml, _, _ = workloads.batch_scene(3, seed=5)
    v, f = workloads.displaced_sphere(8, [0.0, 1.2, 0.0], 1.0, 77)
    path = os.path.join(str(tmp_path), "m.ply")
    workloads.write_ply(path, v, f)
    _, _, a = lm.load_scene(toml + f'\n[[mesh]]\nfile = "{path}"\n'
                                   'material = { type = "pbr", color = [0.7, 0.6, 0.5], roughness = 0.3 }\n')
    _, _, b = lm.load_scene(toml)
    b.add_triangles(v, f, material=b.add_material(workloads._microfacet([0.7, 0.6, 0.5], 0.3)))
    pa, pb = a.prims, b.prims
    assert len(pa) == len(pb)
    tri_a = np.array([list(p.p)[:9] for p in pa if p.kind == _abi.PRIM_TRIANGLE])
    tri_b = np.array([list(p.p)[:9] for p in pb if p.kind == _abi.PRIM_TRIANGLE])
    np.testing.assert_allclose(tri_a, tri_b, rtol=1e-15)


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference example scenes not present")
def test_example_scene_sweep():
    """Every example scene parses: it loads, or it is refused with
    LMN_ERR_UNSUPPORTED naming a feature outside the render path -- never a
    TOML or schema error."""
    files = sorted(glob.glob(os.path.join(REF_SCENES, "*.toml")))
    assert len(files) == 21
    loaded = 0
    for f in files:
        try:
            lm.load_scene(open(f).read(), base_dir=os.path.dirname(f))
            loaded += 1
        except _abi.LuminaraError as e:
            assert e.code == _abi.LMN_ERR_UNSUPPORTED, (f, str(e))
            assert "not on the GPU path" in str(e), (f, str(e))
    assert loaded >= 2  # cornell.toml, prism.toml


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference example scenes not present")
def test_cornell_example_geometry():
    f = os.path.join(REF_SCENES, "cornell.toml")
    cfg, cam, world = lm.load_scene(open(f).read(), base_dir=REF_SCENES)
    assert (cfg.width, cfg.height, cfg.samples_per_pixel) == (600, 600, 256)
    assert len(world.prims) == 18
    assert len(world.lights) >= 1


def test_bad_mesh_leaves_the_description_unchanged():
    # a mesh with an out-of-range vertex or material index is refused before
    # anything is appended: the description keeps its prims (readable, same
    # values), its count and object numbering, and takes a good mesh after
    _, _, world = lm.load_scene(workloads.cornell_box(16, 16, 1, 2, 42))
    before = [(p.kind, p.material, p.object, tuple(p.p)) for p in world.prims]
    n_obj = max(p.object for p in world.prims) + 1
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    big = np.zeros((200000, 3), dtype=np.uint32)  # would force a reallocation
    big[-1] = [0, 1, 3]  # vertex 3 does not exist
    with pytest.raises(lm.LuminaraError):
        world.add_triangles(verts, big, material=0)
    bad_mat = np.zeros((1000, 3), dtype=np.uint32)
    bad_mat[:] = [0, 1, 2]
    with pytest.raises(lm.LuminaraError):
        world.add_triangles(verts, bad_mat, material=10 ** 6)
    after = [(p.kind, p.material, p.object, tuple(p.p)) for p in world.prims]
    assert after == before
    world.add_triangles(verts, np.array([[0, 1, 2]], dtype=np.uint32), material=0)
    prims = world.prims
    assert len(prims) == len(before) + 1
    assert prims[-1].kind == _abi.PRIM_TRIANGLE and prims[-1].object == n_obj
    np.testing.assert_array_equal(np.array(prims[-1].p[:9]).reshape(3, 3), verts)


# ---------------------------------------------------------------- OBJ meshes (src/obj.rs:18-153)

def _obj_scene(tmp_path, text, scale=None, offset=None, name="m.obj"):
    path = tmp_path / name
    path.write_text(text)
    extra = (f"scale = {scale}\n" if scale is not None else "") + (f"offset = {offset}\n" if offset else "")
    toml = (f'[[mesh]]\nfile = "{path}"\n{extra}'
            'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n')
    return lm.load_scene(toml)[2]


@pytest.mark.parametrize("text,n", [
    ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n", 2),                       # test_load_quad_triangulates
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", 1),                                   # test_load_triangle
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n", 1),    # test_load_with_uvs
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\nvn 0 0 1\nvn 0 0 1\n"
     "f 1/1/1 2/2/2 3/3/3\n", 1),                                                  # test_load_full_format
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0 1\nvn 0 0 1\nf 1//1 2//2 3//3\n", 1),  # test_load_with_normals
    ("# comment\nv 0 0 0\nv 1 0 0\nv 0 1 0\n\nf 1 2 3\n", 1),                       # test_obj_with_comments
])
def test_obj_reference_cases(tmp_path, text, n):  # src/obj.rs:328-393
    w = _obj_scene(tmp_path, text)
    assert len(w.prims) == n


def test_obj_scale_offset_and_attributes(tmp_path):  # src/obj.rs:345-356, :114-149
    # scale then offset on positions only; vn made unit; a face has normals
    # (uvs) only when every corner names one
    text = ("v 1 1 1\nv 2 1 1\nv 1 2 1\nv 2 2 1\n"
            "vt 0.25 0.5\nvt 1 0\nvt 0 1\n"
            "vn 0 0 2\nvn 0 3 4\nvn 1 0 0\n"
            "f 1/1/1 2/2/2 3/3/3\n"    # normals and uvs
            "f 2//2 4//1 3//3\n"       # normals only
            "f 2/2 4/3 3\n")           # a corner without uv: plain triangle
    w = _obj_scene(tmp_path, text, scale=2.0, offset="[1.0, 0.0, 0.0]")
    prims = w.prims
    assert [p.kind for p in prims] == [_abi.PRIM_MESH_TRIANGLE, _abi.PRIM_MESH_TRIANGLE, _abi.PRIM_TRIANGLE]
    np.testing.assert_allclose(np.array(prims[0].p[:9]).reshape(3, 3), [[3, 2, 2], [5, 2, 2], [3, 4, 2]])
    d = w.desc_struct
    assert d.n_vertex_attrs == 2
    a0, a1 = d.vertex_attrs[int(prims[0].p[9])], d.vertex_attrs[int(prims[1].p[9])]
    assert a0.has_normals and a0.has_uvs and a1.has_normals and not a1.has_uvs
    np.testing.assert_allclose(np.array(a0.normal), [[0, 0, 1], [0, 0.6, 0.8], [1, 0, 0]])
    np.testing.assert_allclose(np.array(a0.uv), [[0.25, 0.5], [1, 0], [0, 1]])
    np.testing.assert_allclose(np.array(a1.normal), [[0, 0.6, 0.8], [0, 0, 1], [1, 0, 0]])


@pytest.mark.parametrize("text,msg", [
    ("v 0 0\n", "Line 1: vertex needs 3 coordinates"),
    ("v 0 0 0\nvt 1\n", "Line 2: texture coord needs 2 values"),
    ("vn 0 1\n", "Line 1: normal needs 3 components"),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "Line 3: face needs at least 3 vertices"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "Vertex index 4 out of range"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "Vertex index 0 out of range"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -3\n", "Line 4:"),
])
def test_obj_errors(tmp_path, text, msg):
    with pytest.raises(lm.LuminaraError) as e:
        _obj_scene(tmp_path, text)
    assert msg in str(e.value)


def test_obj_out_of_range_attribute_indices_fall_back(tmp_path):
    # normals / uvs that do not exist leave the face without them
    # (obj.rs:114-149: get(n - 1) finds nothing)
    w = _obj_scene(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//2 3//1\n")
    assert [p.kind for p in w.prims] == [_abi.PRIM_TRIANGLE]


def test_c_api_mesh_with_attributes():
    _, _, w = lm.load_scene('[[sphere]]\ncenter = [0.0, -9.0, 0.0]\nradius = 1.0\n'
                            'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }\n')
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=np.float64)
    tris = np.array([[0, 1, 2], [1, 3, 2]], dtype=np.uint32)
    nrm = np.tile(np.array([0.0, 0.0, 1.0]), (2, 3, 1))
    w.add_mesh(verts, tris, normals=nrm, material=0)
    kinds = [p.kind for p in w.prims]
    assert kinds[-2:] == [_abi.PRIM_MESH_TRIANGLE] * 2 and w.desc_struct.n_vertex_attrs == 2
    with pytest.raises(lm.LuminaraError):
        w.add_mesh(verts, np.array([[0, 1, 7]], dtype=np.uint32), normals=nrm[:1])
    assert w.desc_struct.n_vertex_attrs == 2 and len(w.prims) == 3
