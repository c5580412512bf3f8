"""ctypes mirror of include/luminara.h and the loader of liblmn.so.

The library is built in-tree by luminara_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for the render path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# LMN_LIB names another build of the same library (e.g. a tuning experiment)
LIB_PATH = os.environ.get("LMN_LIB") or os.path.join(HERE, "_lib", "liblmn.so")

LMN_OK = 0
LMN_ERR_INVALID = 1
LMN_ERR_PARSE = 2
LMN_ERR_IO = 3
LMN_ERR_UNSUPPORTED = 4
LMN_ERR_DEVICE = 5
LMN_ERR_NOMEM = 6

PRIM_KINDS = ["sphere", "quad", "rect_xy", "rect_xz", "rect_yz", "triangle", "plane", "torus",
              "superellipsoid", "spring", "disk", "cylinder", "cone", "ellipsoid", "hemisphere", "annulus",
              "facet", "wedge_face", "rounded_box", "mobius", "mesh_triangle"]
(PRIM_SPHERE, PRIM_QUAD, PRIM_RECT_XY, PRIM_RECT_XZ, PRIM_RECT_YZ, PRIM_TRIANGLE, PRIM_PLANE,
 PRIM_TORUS, PRIM_SUPERELLIPSOID, PRIM_SPRING, PRIM_DISK, PRIM_CYLINDER, PRIM_CONE,
 PRIM_ELLIPSOID, PRIM_HEMISPHERE, PRIM_ANNULUS, PRIM_FACET, PRIM_WEDGE_FACE, PRIM_ROUNDED_BOX,
 PRIM_MOBIUS, PRIM_MESH_TRIANGLE) = range(21)
MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_EMISSIVE, MAT_MICROFACET = range(5)
LIGHT_SPHERE, LIGHT_RECT, LIGHT_DISK = range(3)
BG_SKY, BG_SOLID, BG_GRADIENT, BG_SUN = range(4)
FILTER_BOX, FILTER_TRIANGLE, FILTER_GAUSSIAN, FILTER_MITCHELL = range(4)
PATH_WAVEFRONT_FP32, PATH_PLAIN_FP64, PATH_PLAIN_FP32 = range(3)
ISECT_EXACT, ISECT_REFERENCE = range(2)
XF_ROTATE_X, XF_ROTATE_Y, XF_ROTATE_Z, XF_TRANSLATE, XF_SCALE, XF_SCALE_XYZ = range(1, 7)

D3 = C.c_double * 3


class Prim(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("material", C.c_uint32), ("transform", C.c_int32),
                ("object", C.c_uint32), ("p", C.c_double * 12)]


class XformOp(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("v", D3)]


class Transform(C.Structure):
    _fields_ = [("n_ops", C.c_uint32), ("pad", C.c_uint32), ("ops", XformOp * 6)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("texture", C.c_int32), ("color", D3), ("fuzz", C.c_double),
                ("ior", C.c_double), ("roughness", C.c_double), ("dispersion", C.c_double),
                ("intensity", C.c_double), ("metallic", C.c_double)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float))]


class Light(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("material", C.c_uint32), ("origin", D3), ("u", D3), ("v", D3),
                ("normal", D3), ("radius", C.c_double), ("emission", D3)]


class Background(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("color0", D3), ("color1", D3),
                ("direction", D3), ("intensity", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("origin", D3), ("lower_left", D3), ("horizontal", D3), ("vertical", D3), ("u", D3),
                ("v", D3), ("lens_radius", C.c_double)]


class RenderConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("max_depth", C.c_uint32), ("seed", C.c_uint64), ("filter", C.c_uint32),
                ("has_crop", C.c_uint32), ("crop", C.c_uint32 * 4), ("exposure", C.c_double),
                ("tone_map", C.c_uint32), ("auto_exposure", C.c_uint32), ("background", Background),
                ("adaptive", C.c_uint32), ("pad", C.c_uint32), ("adaptive_threshold", C.c_double)]


class VertexAttr(C.Structure):
    _fields_ = [("normal", (C.c_double * 3) * 3), ("uv", (C.c_double * 2) * 3), ("has_normals", C.c_uint32),
                ("has_uvs", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("n_prims", C.c_uint32), ("n_transforms", C.c_uint32), ("n_materials", C.c_uint32),
                ("n_textures", C.c_uint32), ("n_lights", C.c_uint32), ("n_objects", C.c_uint32),
                ("prims", C.POINTER(Prim)), ("transforms", C.POINTER(Transform)),
                ("materials", C.POINTER(Material)), ("textures", C.POINTER(Texture)),
                ("lights", C.POINTER(Light)), ("camera", Camera), ("config", RenderConfig),
                ("n_vertex_attrs", C.c_uint64), ("vertex_attrs", C.POINTER(VertexAttr))]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("seed", C.c_uint64), ("filter", C.c_uint32),
                ("has_crop", C.c_uint32), ("crop", C.c_uint32 * 4), ("path", C.c_uint32),
                ("intersect_mode", C.c_uint32), ("scene_key", C.c_uint32), ("pool_paths", C.c_uint32),
                ("block_size", C.c_uint32), ("stack_lds", C.c_uint32), ("flags", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("stream", C.c_void_p), ("adaptive_threshold", C.c_double),
                ("time_limit", C.c_double)]

FLAG_STAGE_TIMING = 1


class Stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("render_ms", C.c_double), ("iterations", C.c_uint32),
                ("extend_launches", C.c_uint32), ("extend_ms", C.c_double), ("shade_ms", C.c_double),
                ("shadow_ms", C.c_double), ("regen_ms", C.c_double), ("march_ms", C.c_double),
                ("samples_traced", C.c_uint64), ("rounds", C.c_uint32), ("expired_rows", C.c_uint32)]


class Ray(C.Structure):
    _fields_ = [("origin", D3), ("direction", D3), ("t_min", C.c_double), ("t_max", C.c_double)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_double), ("prim", C.c_int64), ("u", C.c_double), ("v", C.c_double),
                ("normal", D3), ("front_face", C.c_int32), ("pad", C.c_int32)]


class TraceCounts(C.Structure):
    _fields_ = [("node_fetches", C.c_uint64), ("prim_tests", C.c_uint64), ("tri_tests", C.c_uint64),
                ("rays", C.c_uint64), ("node_bytes", C.c_uint32), ("prim_bytes", C.c_uint32), ("hits", C.c_uint64)]


FLAG_NO_GRAPH = 2
FLAG_ADAPTIVE = 4
FLAG_PLAIN_WAVES_SHIFT = 8  # bits 8-10: waves per SIMD of the plain FP64 kernel build (2, 3, 4)


class Tuning(C.Structure):
    _fields_ = [("pool_paths", C.c_uint32), ("block_size", C.c_uint32), ("stack_lds", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("probe_ms", C.c_double), ("default_ms", C.c_double),
                ("probe_paths", C.c_uint64), ("candidates", C.c_uint32), ("pad", C.c_uint32)]


ADJOINT_MAX_DEPTH = 50


class AdjointSpec(C.Structure):
    _fields_ = [("texture", C.c_int32), ("emission_material", C.c_int32), ("flags", C.c_uint32),
                ("pad", C.c_uint32)]


class Grads(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("roughness", C.c_void_p), ("emission", C.c_double),
                ("paths", C.c_uint64), ("adjoint_ms", C.c_double)]


TONE_ACES, TONE_REINHARD, TONE_FILMIC, TONE_NONE = 0, 1, 2, 3


class PostParams(C.Structure):
    _fields_ = [("exposure", C.c_double), ("tone_map", C.c_uint32), ("auto_exposure", C.c_uint32),
                ("depth_fog", C.c_double), ("fog_color", D3)]


class Image(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("hdr", C.c_void_p), ("rgbe", C.c_void_p), ("depth", C.c_void_p),
                ("normal", C.c_void_p), ("albedo", C.c_void_p), ("exposure", C.c_double),
                ("log_average", C.c_double), ("post_ms", C.c_double)]


# Every symbol include/luminara.h declares, with its ctypes signature.
SIGNATURES = {
    "lmn_last_error": (C.c_char_p, []),
    "lmn_load_scene": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.POINTER(SceneDesc))]),
    "lmn_desc_free": (None, [C.POINTER(SceneDesc)]),
    "lmn_desc_set_texture": (C.c_int, [C.POINTER(SceneDesc), C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float)]),
    "lmn_desc_add_texture": (C.c_int, [C.POINTER(SceneDesc), C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                       C.POINTER(C.c_uint32)]),
    "lmn_desc_set_material": (C.c_int, [C.POINTER(SceneDesc), C.c_uint32, C.POINTER(Material)]),
    "lmn_desc_add_material": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(Material), C.POINTER(C.c_uint32)]),
    "lmn_desc_add_triangles": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_double), C.c_uint64,
                                         C.POINTER(C.c_uint32), C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32)]),
    "lmn_desc_add_mesh": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_uint32),
                                    C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "lmn_scene_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "lmn_scene_destroy": (None, [C.c_void_p]),
    "lmn_scene_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "lmn_render": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.POINTER(C.c_float), C.POINTER(Stats)]),
    "lmn_render_device": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.POINTER(Stats)]),
    "lmn_render_batch": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(RenderParams),
                                   C.POINTER(C.POINTER(C.c_float)), C.POINTER(Stats)]),
    "lmn_render_batch_device": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(RenderParams),
                                          C.POINTER(C.c_void_p), C.POINTER(Stats)]),
    "lmn_trace_rays": (C.c_int, [C.c_void_p, C.POINTER(Ray), C.c_size_t, C.c_uint32, C.c_uint32,
                                 C.POINTER(Hit)]),
    "lmn_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "lmn_profile_extend": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(RenderParams),
                                     C.POINTER(C.c_void_p), C.POINTER(TraceCounts)]),
    "lmn_autotune": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(RenderParams), C.POINTER(C.c_void_p),
                               C.POINTER(Tuning)]),
    "lmn_adjoint": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.POINTER(AdjointSpec), C.POINTER(C.c_float),
                              C.POINTER(Grads)]),
    "lmn_adjoint_device": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.POINTER(AdjointSpec), C.c_void_p,
                                     C.POINTER(Grads)]),
    "lmn_aov": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.c_void_p, C.c_void_p]),
    "lmn_aov_device": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.c_void_p, C.c_void_p]),
    "lmn_post": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(PostParams), C.c_void_p, C.c_void_p,
                           C.c_void_p, C.POINTER(C.c_double)]),
    "lmn_post_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(PostParams), C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "lmn_post_device_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(PostParams), C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "lmn_encode_rgbe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "lmn_encode_rgbe_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "lmn_render_image": (C.c_int, [C.c_void_p, C.POINTER(RenderParams), C.POINTER(PostParams), C.POINTER(Image),
                                   C.POINTER(Stats)]),
}

_lib = None


class LuminaraError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"luminara error {code}: {msg}")
        self.code = code


def lib():
    """The loaded liblmn.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python luminara_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != LMN_OK:
        raise LuminaraError(rc, lib().lmn_last_error().decode())
    return rc
