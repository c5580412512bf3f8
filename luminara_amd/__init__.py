"""MI355X-native render hot path of luminara (HIP kernels behind a C ABI).

Host-side mirror of the reference's render entry:

    config, camera, world = load_scene(toml_text)        # scene::load_scene (src/scene.rs:846)
    result = render(config, camera, world, world.lights)  # render::render (src/render.rs:1557)
    result.rows            # H x W x 3 linear HDR, row 0 = top (src/render.rs:1598-1731)
    result.total_rays      # primary samples (src/render.rs:1704)
    result.render_time_secs

All work runs in liblmn.so (HIP, gfx950); nothing here computes pixels.
"""
import ctypes as C
import dataclasses
import os
from typing import Optional, Tuple

import numpy as np

from . import _abi
from ._abi import LuminaraError, check, lib

FILTERS = {"box": _abi.FILTER_BOX, "triangle": _abi.FILTER_TRIANGLE, "gaussian": _abi.FILTER_GAUSSIAN,
           "mitchell": _abi.FILTER_MITCHELL}
PATHS = {"wavefront": _abi.PATH_WAVEFRONT_FP32, "plain64": _abi.PATH_PLAIN_FP64,
         "plain32": _abi.PATH_PLAIN_FP32}


@dataclasses.dataclass
class RenderConfig:
    """Hot-path subset of the reference's RenderConfig (src/render.rs:239-521)."""
    width: int = 800
    height: int = 450
    samples_per_pixel: int = 100
    max_depth: int = 50
    seed: int = 31337
    pixel_filter: int = _abi.FILTER_BOX
    crop: Optional[Tuple[int, int, int, int]] = None
    exposure: float = 1.0
    tone_map: int = 0
    auto_exposure: bool = False
    adaptive: bool = False
    adaptive_threshold: float = 0.03
    time_limit: float = 0.0
    depth_fog: float = 0.0
    depth_fog_color: Tuple[float, float, float] = (0.8, 0.85, 0.9)


class SceneWorld:
    """A loaded scene: the flattened descriptor plus (lazily) its device copy."""

    def __init__(self, desc_ptr):
        self._desc = desc_ptr
        self._scene = None

    @property
    def desc(self):
        return self._desc

    @property
    def desc_struct(self) -> _abi.SceneDesc:
        return self._desc.contents

    @property
    def lights(self):
        d = self.desc_struct
        return [d.lights[i] for i in range(d.n_lights)]

    @property
    def prims(self):
        d = self.desc_struct
        return [d.prims[i] for i in range(d.n_prims)]

    def materials(self):
        d = self.desc_struct
        return [d.materials[i] for i in range(d.n_materials)]

    def set_material(self, index, mat: _abi.Material):
        check(lib().lmn_desc_set_material(self._desc, index, C.byref(mat)))
        self.invalidate()

    def add_texture(self, rgb: np.ndarray) -> int:
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w, _ = rgb.shape
        idx = C.c_uint32()
        check(lib().lmn_desc_add_texture(self._desc, w, h, rgb.ctypes.data_as(C.POINTER(C.c_float)),
                                         C.byref(idx)))
        self.invalidate()
        return idx.value

    def add_material(self, mat: _abi.Material) -> int:
        idx = C.c_uint32()
        check(lib().lmn_desc_add_material(self._desc, C.byref(mat), C.byref(idx)))
        self.invalidate()
        return idx.value

    def add_triangles(self, verts: np.ndarray, tris: np.ndarray, material: int = 0, materials=None):
        """Appends a triangle mesh (one prim and world object per face)."""
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.uint32)
        check(lib().lmn_desc_add_triangles(
            self._desc, v.ctypes.data_as(C.POINTER(C.c_double)), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)),
            len(t), material, None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint32))))
        self.invalidate()

    def add_mesh(self, verts: np.ndarray, tris: np.ndarray, normals=None, uvs=None, material: int = 0,
                 materials=None):
        """Appends a triangle mesh with per-corner normals (n_tris x 3 x 3) and/or
        texture coordinates (n_tris x 3 x 2): MESH_TRIANGLE prims."""
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(len(t), 9)
        uv = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float64).reshape(len(t), 6)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.uint32)
        check(lib().lmn_desc_add_mesh(
            self._desc, v.ctypes.data_as(C.POINTER(C.c_double)), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)),
            len(t), None if nrm is None else nrm.ctypes.data, None if uv is None else uv.ctypes.data, material,
            None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint32))))
        self.invalidate()

    def set_texture(self, index, rgb: np.ndarray):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w, _ = rgb.shape
        check(lib().lmn_desc_set_texture(self._desc, index, w, h, rgb.ctypes.data_as(C.POINTER(C.c_float))))
        self.invalidate()

    def invalidate(self):
        if self._scene is not None:
            lib().lmn_scene_destroy(self._scene)
            self._scene = None

    def device_scene(self):
        """Flattens, builds the BVH and uploads to the current HIP device."""
        if self._scene is None:
            h = C.c_void_p()
            check(lib().lmn_scene_create(self._desc, C.byref(h)))
            self._scene = h
        return self._scene

    def bvh_info(self):
        n_nodes, n_leaves, dev_bytes = C.c_uint64(), C.c_uint64(), C.c_uint64()
        depth = C.c_uint32()
        check(lib().lmn_scene_info(self.device_scene(), C.byref(n_nodes), C.byref(n_leaves), C.byref(depth),
                                   C.byref(dev_bytes)))
        return {"nodes": n_nodes.value, "leaves": n_leaves.value, "depth": depth.value,
                "device_bytes": dev_bytes.value}

    def __del__(self):
        try:
            self.invalidate()
            if self._desc is not None:
                lib().lmn_desc_free(self._desc)
                self._desc = None
        except Exception:
            pass


def load_scene(toml_text: str, base_dir: Optional[str] = None):
    """scene::load_scene: returns (RenderConfig, camera, SceneWorld).

    Raises LuminaraError (code LMN_ERR_PARSE / LMN_ERR_IO / LMN_ERR_UNSUPPORTED)
    where the reference returns Err(String).
    """
    out = C.POINTER(_abi.SceneDesc)()
    check(lib().lmn_load_scene(toml_text.encode(), base_dir.encode() if base_dir else None, C.byref(out)))
    world = SceneWorld(out)
    c = out.contents.config
    cfg = RenderConfig(width=c.width, height=c.height, samples_per_pixel=c.samples, max_depth=c.max_depth,
                       seed=c.seed, pixel_filter=c.filter, exposure=c.exposure, tone_map=c.tone_map,
                       auto_exposure=bool(c.auto_exposure), adaptive=bool(c.adaptive),
                       adaptive_threshold=c.adaptive_threshold)
    return cfg, out.contents.camera, world


def make_params(config: RenderConfig, path="wavefront", scene_key=0, intersect_mode=_abi.ISECT_EXACT,
                pool_paths=0, stack_lds=0, stream=None, flags=0, block_size=0, grid_blocks=0) -> _abi.RenderParams:
    p = _abi.RenderParams()
    p.width, p.height = config.width, config.height
    p.spp, p.max_depth = config.samples_per_pixel, config.max_depth
    p.seed = config.seed
    p.filter = config.pixel_filter
    if config.crop is not None:
        p.has_crop = 1
        for i, v in enumerate(config.crop):
            p.crop[i] = v
    p.path = PATHS[path] if isinstance(path, str) else path
    p.intersect_mode = intersect_mode
    p.scene_key = scene_key
    p.pool_paths = pool_paths
    p.stack_lds = stack_lds
    p.block_size = block_size
    p.grid_blocks = grid_blocks
    p.flags = flags | (_abi.FLAG_ADAPTIVE if config.adaptive else 0)
    p.stream = stream
    p.adaptive_threshold = config.adaptive_threshold
    p.time_limit = config.time_limit
    return p


def output_shape(config: RenderConfig):
    w, h = max(config.width, 1), max(config.height, 1)
    if config.crop is not None:
        cx, cy, cw, ch = config.crop
        return min(ch, h), min(cw, w)
    return h, w


@dataclasses.dataclass
class RenderResult:
    """RenderResult (src/render.rs:1540-1554), HDR part."""
    rows: np.ndarray
    total_rays: int
    render_time_secs: float
    stats: dict


def render(config: RenderConfig, camera, world: SceneWorld, lights=None, path="wavefront", **kw) -> RenderResult:
    """render::render's pixel loop on the GPU.  `camera` and `lights` are the
    ones load_scene produced (the descriptor holds them); CLI overrides of
    width/height do not rebuild the camera (src/main.rs:205-210)."""
    del camera, lights  # carried by the scene descriptor
    scene = world.device_scene()
    params = make_params(config, path=path, **kw)
    h, w = output_shape(config)
    out = np.empty((h, w, 3), dtype=np.float32)
    st = _abi.Stats()
    check(lib().lmn_render(scene, C.byref(params), out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in _abi.Stats._fields_}
    return RenderResult(rows=out, total_rays=st.paths, render_time_secs=st.render_ms / 1000.0, stats=stats)


def render_device(world: SceneWorld, params: _abi.RenderParams, out_ptr: int) -> dict:
    """Renders into a caller-owned device buffer (e.g. a torch CUDA tensor)."""
    st = _abi.Stats()
    check(lib().lmn_render_device(world.device_scene(), C.byref(params), C.c_void_p(out_ptr), C.byref(st)))
    return {k: getattr(st, k) for k, _ in _abi.Stats._fields_}


def render_batch(configs, worlds, path="wavefront", **kw):
    """Renders several scenes in one wavefront (lmn_render_batch); returns
    the list of framebuffers and the batch statistics."""
    n = len(worlds)
    scenes = (C.c_void_p * n)(*[w.device_scene().value for w in worlds])
    params = (_abi.RenderParams * n)(*[make_params(c, path=path, scene_key=i, **kw) for i, c in enumerate(configs)])
    outs = [np.empty(output_shape(c) + (3,), dtype=np.float32) for c in configs]
    ptrs = (C.POINTER(C.c_float) * n)(*[o.ctypes.data_as(C.POINTER(C.c_float)) for o in outs])
    st = _abi.Stats()
    check(lib().lmn_render_batch(scenes, n, params, ptrs, C.byref(st)))
    return outs, {k: getattr(st, k) for k, _ in _abi.Stats._fields_}


def render_batch_device(worlds, params_list, out_ptrs) -> dict:
    n = len(worlds)
    scenes = (C.c_void_p * n)(*[w.device_scene().value for w in worlds])
    params = (_abi.RenderParams * n)(*params_list)
    ptrs = (C.c_void_p * n)(*out_ptrs)
    st = _abi.Stats()
    check(lib().lmn_render_batch_device(scenes, n, params, ptrs, C.byref(st)))
    return {k: getattr(st, k) for k, _ in _abi.Stats._fields_}


def profile_extend(worlds, params_list, out_ptrs) -> dict:
    """Instrumented wavefront render of a batch: BVH node fetches and
    primitive tests of the extend stage."""
    n = len(worlds)
    scenes = (C.c_void_p * n)(*[w.device_scene().value for w in worlds])
    params = (_abi.RenderParams * n)(*params_list)
    ptrs = (C.c_void_p * n)(*out_ptrs)
    tc = _abi.TraceCounts()
    check(lib().lmn_profile_extend(scenes, n, params, ptrs, C.byref(tc)))
    return {k: getattr(tc, k) for k, _ in _abi.TraceCounts._fields_}


def autotune(worlds, params_list, out_ptrs=None) -> dict:
    """lmn_autotune: picks the wavefront launch configuration for this batch
    (probe renders into out_ptrs, or into library buffers when None), caches
    it, and returns it."""
    n = len(worlds)
    scenes = (C.c_void_p * n)(*[w.device_scene().value for w in worlds])
    params = (_abi.RenderParams * n)(*params_list)
    ptrs = (C.c_void_p * n)(*out_ptrs) if out_ptrs is not None else None
    t = _abi.Tuning()
    check(lib().lmn_autotune(scenes, n, params, ptrs, C.byref(t)))
    return {k: getattr(t, k) for k, _ in _abi.Tuning._fields_ if k != "pad"}


def rays_to_struct(rays: np.ndarray):
    """rays: (n, 8) float64 = origin.xyz, direction.xyz, t_min, t_max."""
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    assert rays.ndim == 2 and rays.shape[1] == 8
    arr = (_abi.Ray * len(rays)).from_buffer_copy(rays.tobytes())
    return arr


def hits_to_numpy(hits, n):
    buf = np.frombuffer(bytes(hits), dtype=np.dtype([("t", "f8"), ("prim", "i8"), ("u", "f8"), ("v", "f8"),
                                                     ("normal", "f8", 3), ("front_face", "i4"), ("pad", "i4")]),
                        count=n)
    return buf


def trace_rays(world: SceneWorld, rays: np.ndarray, path="wavefront", intersect_mode=_abi.ISECT_EXACT):
    """Closest hit per ray on the GPU (SceneWorld::hit, src/scene.rs:817)."""
    n = len(rays)
    r = rays_to_struct(rays)
    hits = (_abi.Hit * max(n, 1))()
    check(lib().lmn_trace_rays(world.device_scene(), r, n, PATHS[path] if isinstance(path, str) else path,
                               intersect_mode, hits))
    return hits_to_numpy(hits, n)


@dataclasses.dataclass
class Gradients:
    """Output of the path-replay adjoint."""
    texels: Optional[np.ndarray]     # (h, w, 3) float32, d loss / d texel of the parameter texture
    roughness: np.ndarray            # (n_materials,) float32, d loss / d roughness (microfacet materials)
    emission: float                  # d loss / d intensity of the parameter emissive material
    paths: int
    adjoint_ms: float


def adjoint(config: RenderConfig, world: SceneWorld, dl_dimage: np.ndarray, texture=-1, emission_material=-1,
            path="wavefront", **kw) -> Gradients:
    """Gradients of an image loss (given dL/dI for the image `render` returns
    with the same arguments) by the path-replay adjoint (lmn_adjoint)."""
    scene = world.device_scene()
    params = make_params(config, path=path, **kw)
    h, w = output_shape(config)
    g = np.ascontiguousarray(dl_dimage, dtype=np.float32).reshape(h, w, 3)
    spec = _abi.AdjointSpec(texture, emission_material, 0, 0)
    n_mat = world.desc_struct.n_materials
    rough = np.zeros(n_mat, dtype=np.float32)
    tex = None
    if texture >= 0:
        if texture >= world.desc_struct.n_textures:
            raise ValueError(f"texture {texture} out of range")
        t = world.desc_struct.textures[texture]
        tex = np.zeros((t.height, t.width, 3), dtype=np.float32)
    gr = _abi.Grads()
    gr.texels = tex.ctypes.data if tex is not None else None
    gr.roughness = rough.ctypes.data if n_mat else None
    check(lib().lmn_adjoint(scene, C.byref(params), C.byref(spec), g.ctypes.data_as(C.POINTER(C.c_float)),
                            C.byref(gr)))
    return Gradients(tex, rough, gr.emission, gr.paths, gr.adjoint_ms)


def adjoint_device(world: SceneWorld, params: _abi.RenderParams, spec: _abi.AdjointSpec, dl_ptr: int,
                   texels_ptr: Optional[int], rough_ptr: Optional[int]) -> dict:
    """lmn_adjoint_device on caller-owned device buffers (e.g. torch tensors)."""
    gr = _abi.Grads()
    gr.texels = texels_ptr
    gr.roughness = rough_ptr
    check(lib().lmn_adjoint_device(world.device_scene(), C.byref(params), C.byref(spec), C.c_void_p(dl_ptr),
                                   C.byref(gr)))
    return {"emission": gr.emission, "paths": gr.paths, "adjoint_ms": gr.adjoint_ms}


def post_params(config: RenderConfig) -> _abi.PostParams:
    pp = _abi.PostParams()
    pp.exposure = config.exposure
    pp.tone_map = config.tone_map
    pp.auto_exposure = 1 if config.auto_exposure else 0
    pp.depth_fog = config.depth_fog
    for k in range(3):
        pp.fog_color[k] = config.depth_fog_color[k]
    return pp


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def aov(config: RenderConfig, world: SceneWorld, depth=True, normal=True, albedo=True, path="wavefront",
        **kw):
    """AOV passes of one frame (src/render.rs:3508-3584): (depth, normal,
    albedo) arrays of the crop, None for passes not asked for.  depth is +inf
    where the pixel's ray escapes."""
    h, w = output_shape(config)
    d = np.empty((h, w), np.float32) if depth else None
    n = np.empty((h, w, 3), np.uint8) if normal else None
    a = np.empty((h, w, 3), np.uint8) if albedo else None
    params = make_params(config, path=path, **kw)
    check(lib().lmn_aov(world.device_scene(), C.byref(params), _ptr(d), _ptr(n), _ptr(a)))
    return d, n, a


def post(hdr: np.ndarray, config: RenderConfig, fog_depth=None, want_hdr=False):
    """Default post chain of a (H, W, 3) float32 framebuffer on the GPU:
    returns (rgba (H, W, 4) uint8, exposed HDR or None, exposure used)."""
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    h, w = hdr.shape[:2]
    rgba = np.empty((h, w, 4), np.uint8)
    hx = np.empty((h, w, 3), np.float32) if want_hdr else None
    fog = None if fog_depth is None else np.ascontiguousarray(fog_depth, dtype=np.float32)
    ex = C.c_double()
    pp = post_params(config)
    check(lib().lmn_post(_ptr(hdr), w, h, C.byref(pp), _ptr(fog), _ptr(rgba), _ptr(hx), C.byref(ex)))
    return rgba, hx, ex.value


def encode_rgbe(rgb: np.ndarray) -> np.ndarray:
    """Radiance RGBE pixels of (..., 3) float32 data on the GPU."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    n = rgb.size // 3
    out = np.empty(rgb.shape[:-1] + (4,), np.uint8)
    check(lib().lmn_encode_rgbe(_ptr(rgb), n, _ptr(out)))
    return out


def aov_device(world: SceneWorld, params: _abi.RenderParams, depth_ptr=0, normal_ptr=0, albedo_ptr=0):
    """lmn_aov_device into caller-owned device buffers (0 = skip), ordered on
    params.stream."""
    check(lib().lmn_aov_device(world.device_scene(), C.byref(params), C.c_void_p(depth_ptr or None),
                               C.c_void_p(normal_ptr or None), C.c_void_p(albedo_ptr or None)))


def post_device(hdr_ptr: int, width: int, height: int, config: RenderConfig, fog_ptr=0, rgba_ptr=0, hdr_out_ptr=0,
                stream=None) -> float:
    """lmn_post_device on device buffers, ordered on `stream` (a HIP stream
    handle, None = the library's); returns the exposure applied."""
    ex = C.c_double()
    pp = post_params(config)
    check(lib().lmn_post_device(C.c_void_p(hdr_ptr), width, height, C.byref(pp), C.c_void_p(fog_ptr or None),
                                C.c_void_p(rgba_ptr or None), C.c_void_p(hdr_out_ptr or None), C.byref(ex),
                                C.c_void_p(stream) if stream else None))
    return ex.value


def encode_rgbe_device(rgb_ptr: int, n_pixels: int, out_ptr: int, stream=None):
    check(lib().lmn_encode_rgbe_device(C.c_void_p(rgb_ptr), n_pixels, C.c_void_p(out_ptr),
                                       C.c_void_p(stream) if stream else None))


@dataclasses.dataclass
class ImageResult:
    """RenderResult (src/render.rs:1540-1554) with the default post chain."""
    pixels: np.ndarray
    hdr_data: Optional[np.ndarray]
    rgbe: Optional[np.ndarray]
    depth_pass: Optional[np.ndarray]
    normal_pass: Optional[np.ndarray]
    albedo_pass: Optional[np.ndarray]
    exposure: float
    total_rays: int
    render_time_secs: float
    post_ms: float


def render_image(config: RenderConfig, world: SceneWorld, save_hdr=False, save_depth=False, save_normals=False,
                 save_albedo=False, path="wavefront", **kw) -> ImageResult:
    """lmn_render_image: the frame, its AOV ray, post chain and RGBE pixels on
    the GPU; only the outputs asked for come back."""
    h, w = output_shape(config)
    rgba = np.empty((h, w, 4), np.uint8)
    hx = np.empty((h, w, 3), np.float32) if save_hdr else None
    rgbe = np.empty((h, w, 4), np.uint8) if save_hdr else None
    d = np.empty((h, w), np.float32) if save_depth else None
    n = np.empty((h, w, 3), np.uint8) if save_normals else None
    a = np.empty((h, w, 3), np.uint8) if save_albedo else None
    img = _abi.Image(rgba=rgba.ctypes.data, hdr=None if hx is None else hx.ctypes.data,
                     rgbe=None if rgbe is None else rgbe.ctypes.data, depth=None if d is None else d.ctypes.data,
                     normal=None if n is None else n.ctypes.data, albedo=None if a is None else a.ctypes.data)
    params = make_params(config, path=path, **kw)
    pp = post_params(config)
    st = _abi.Stats()
    check(lib().lmn_render_image(world.device_scene(), C.byref(params), C.byref(pp), C.byref(img), C.byref(st)))
    return ImageResult(pixels=rgba, hdr_data=hx, rgbe=rgbe, depth_pass=d, normal_pass=n, albedo_pass=a,
                       exposure=img.exposure, total_rays=st.paths, render_time_secs=st.render_ms / 1000.0,
                       post_ms=img.post_ms)


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int()
    hbm = C.c_double()
    check(lib().lmn_device_info(name, 256, C.byref(ncu), C.byref(hbm)))
    return {"name": name.value.decode(), "cu": ncu.value, "hbm_bytes": hbm.value}


__all__ = ["RenderConfig", "SceneWorld", "RenderResult", "ImageResult", "LuminaraError", "load_scene", "render",
           "render_device", "render_image", "aov", "post", "encode_rgbe", "trace_rays", "device_info",
           "make_params", "profile_extend"]
