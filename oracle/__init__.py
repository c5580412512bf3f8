"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.  The
product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

RNG_COUNTER = 0
RNG_REFERENCE = 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class OracleParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("seed", C.c_uint64), ("filter", C.c_uint32), ("has_crop", C.c_uint32), ("crop", C.c_uint32 * 4),
                ("rng_mode", C.c_uint32), ("scene_key", C.c_uint32), ("threads", C.c_uint32),
                ("exact_marched", C.c_uint32), ("row_begin", C.c_uint32), ("row_end", C.c_uint32),
                ("row_step", C.c_uint32), ("adaptive", C.c_uint32), ("adaptive_threshold", C.c_double),
                ("time_limit", C.c_double)]


class OracleStats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("render_secs", C.c_double), ("threads", C.c_uint32), ("pad", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        from luminara_amd import _abi
        L = C.CDLL(LIB_PATH)
        L.oracle_trace.restype = C.c_int
        L.oracle_trace.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(_abi.Ray), C.c_size_t, C.c_uint32,
                                   C.POINTER(_abi.Hit)]
        L.oracle_render.restype = C.c_int
        L.oracle_render.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(OracleParams),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(OracleStats)]
        L.oracle_scatter.restype = C.c_int
        L.oracle_scatter.argtypes = [C.POINTER(_abi.SceneDesc), C.c_uint32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_uint64, C.c_uint32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
        L.oracle_scatter_batch.restype = C.c_int
        L.oracle_scatter_batch.argtypes = [C.POINTER(_abi.SceneDesc), C.c_uint32, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
        L.oracle_xoshiro.restype = C.c_int
        L.oracle_xoshiro.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_last_error.restype = C.c_char_p
        D3 = C.POINTER(C.c_double)
        L.oracle_nee.restype = C.c_int
        L.oracle_nee.argtypes = [C.POINTER(_abi.SceneDesc), D3, D3, D3, C.c_uint64, C.c_uint32, D3,
                                 C.POINTER(C.c_uint64)]
        L.oracle_camera_ray.restype = C.c_int
        L.oracle_camera_ray.argtypes = [C.POINTER(_abi.SceneDesc), C.c_double, C.c_double, C.c_uint64, D3, D3]
        L.oracle_emitted.restype = C.c_int
        L.oracle_emitted.argtypes = [C.POINTER(_abi.SceneDesc), C.c_uint32, C.c_double, C.c_double, D3]
        L.oracle_background.restype = C.c_int
        L.oracle_background.argtypes = [C.POINTER(_abi.SceneDesc), D3, D3]
        L.oracle_filter_weight.restype = C.c_double
        L.oracle_filter_weight.argtypes = [C.c_uint32, C.c_double, C.c_double]
        L.oracle_adjoint.restype = C.c_int
        L.oracle_adjoint.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(OracleParams), C.c_int32, C.c_int32,
                                     D3, D3, D3, D3]
        L.oracle_reflectance.restype = C.c_double
        L.oracle_reflectance.argtypes = [C.c_double, C.c_double]
        V = C.c_void_p
        L.oracle_aov.restype = C.c_int
        L.oracle_aov.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(OracleParams), V, V, V]
        L.oracle_post.restype = C.c_int
        L.oracle_post.argtypes = [V, C.c_uint64, C.c_double, C.c_uint32, C.c_uint32, V, C.c_double, D3, V, V,
                                  C.POINTER(C.c_double)]
        L.oracle_log_average_luminance.restype = C.c_double
        L.oracle_log_average_luminance.argtypes = [V, C.c_uint64]
        L.oracle_tone.restype = C.c_double
        L.oracle_tone.argtypes = [C.c_uint32, C.c_double]
        L.oracle_srgb8.restype = C.c_uint32
        L.oracle_srgb8.argtypes = [C.c_double]
        L.oracle_rgbe.restype = None
        L.oracle_rgbe.argtypes = [V, C.c_uint64, V]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().oracle_last_error().decode()}")


def trace(world, rays: np.ndarray, exact=True):
    from luminara_amd import hits_to_numpy, rays_to_struct, _abi
    n = len(rays)
    r = rays_to_struct(rays)
    hits = (_abi.Hit * max(n, 1))()
    _check(lib().oracle_trace(world.desc, r, n, 1 if exact else 0, hits))
    return hits_to_numpy(hits, n)


def render(world, config, rng_mode=RNG_COUNTER, threads=0, scene_key=0, exact=True, rows=None,
           want_var=True, row_step=1):
    """Pixel loop on the CPU.  Returns (mean HxWx3, var HxWx3 or None, stats);
    `rows` = (begin, end) restricts it to every `row_step`-th row of a band."""
    p = OracleParams()
    p.width, p.height = config.width, config.height
    p.spp, p.max_depth, p.seed = config.samples_per_pixel, config.max_depth, config.seed
    p.filter = config.pixel_filter
    fw, fh = max(config.width, 1), max(config.height, 1)
    cw, ch = fw, fh
    if config.crop is not None:
        p.has_crop = 1
        for i, v in enumerate(config.crop):
            p.crop[i] = v
        cw, ch = min(config.crop[2], fw), min(config.crop[3], fh)
    p.rng_mode, p.scene_key, p.threads = rng_mode, scene_key, threads
    p.exact_marched = 1 if exact else 0
    p.adaptive = 1 if getattr(config, "adaptive", False) else 0
    p.adaptive_threshold = getattr(config, "adaptive_threshold", 0.03)
    p.time_limit = getattr(config, "time_limit", 0.0)
    rb, re = rows if rows is not None else (0, ch)
    re = min(re, ch)
    p.row_begin, p.row_end, p.row_step = rb, re, row_step
    nrows = max(0, (re - rb + row_step - 1) // row_step)
    mean = np.zeros((nrows, cw, 3), dtype=np.float64)
    var = np.zeros((nrows, cw, 3), dtype=np.float64) if want_var else None
    st = OracleStats()
    _check(lib().oracle_render(world.desc, C.byref(p), mean.ctypes.data_as(C.POINTER(C.c_double)),
                               var.ctypes.data_as(C.POINTER(C.c_double)) if want_var else None, C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in OracleStats._fields_ if k != "pad"}
    return mean, var, stats


def scatter(world, material, ray_dir, point, normal, front_face, t, u, v, key, counter=0):
    D = C.c_double * 3
    od, oa = D(), D()
    oc = C.c_uint32()
    ok = lib().oracle_scatter(world.desc, material, D(*ray_dir), D(*point), D(*normal), int(front_face), t, u, v,
                              key, counter, od, oa, C.byref(oc))
    if ok < 0:
        _check(ok)
    return bool(ok), np.array(od[:]), np.array(oa[:]), oc.value


def scatter_batch(world, material, ray_dir, point, normal, front_face, n, t=1.0, u=0.5, v=0.5, key0=0):
    """n scatters at one point (sample k from stream key0 + k): (ok bool[n],
    directions [n, 3], attenuations [n, 3])."""
    D = C.c_double * 3
    dirs, atts = np.zeros((n, 3)), np.zeros((n, 3))
    ok = np.zeros(n, dtype=np.uint8)
    _check(lib().oracle_scatter_batch(world.desc, material, D(*ray_dir), D(*point), D(*normal), int(front_face), t,
                                      u, v, key0, n, dirs.ctypes.data, atts.ctypes.data, ok.ctypes.data))
    return ok.astype(bool), dirs, atts


def xoshiro(n, state=None, seed=0):
    """The reference stream's generator: n u64 outputs of xoshiro256++ from a
    raw state (4 u64) or from seed_from_u64(seed), and their f64 draws."""
    out = np.zeros(n, dtype=np.uint64)
    f = np.zeros(n)
    st = None if state is None else np.ascontiguousarray(state, dtype=np.uint64)
    _check(lib().oracle_xoshiro(None if st is None else st.ctypes.data, seed, n, out.ctypes.data, f.ctypes.data))
    return out, f


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def nee(world, point, normal, albedo, key=42, n=1000):
    """Mean of n NEE samples at a fixed surface point, and the shadow-ray count."""
    out, sh = (C.c_double * 3)(), C.c_uint64()
    _check(lib().oracle_nee(world.desc, _d3(point), _d3(normal), _d3(albedo), key, n, out, C.byref(sh)))
    return np.array(out[:]), sh.value


def camera_ray(world, s, t, key=0):
    o, d = (C.c_double * 3)(), (C.c_double * 3)()
    _check(lib().oracle_camera_ray(world.desc, s, t, key, o, d))
    return np.array(o[:]), np.array(d[:])


def emitted(world, material, u=0.5, v=0.5):
    out = (C.c_double * 3)()
    _check(lib().oracle_emitted(world.desc, material, u, v, out))
    return np.array(out[:])


def background(world, direction):
    out = (C.c_double * 3)()
    _check(lib().oracle_background(world.desc, _d3(direction), out))
    return np.array(out[:])


def filter_weight(kind, dx, dy):
    return lib().oracle_filter_weight(kind, dx, dy)


def reflectance(cosine, ratio):
    return lib().oracle_reflectance(cosine, ratio)


def _params(config, rng_mode=RNG_COUNTER, threads=0, scene_key=0, exact=True):
    p = OracleParams()
    p.width, p.height = config.width, config.height
    p.spp, p.max_depth, p.seed = config.samples_per_pixel, config.max_depth, config.seed
    p.filter = config.pixel_filter
    if config.crop is not None:
        p.has_crop = 1
        for i, v in enumerate(config.crop):
            p.crop[i] = v
    p.rng_mode, p.scene_key, p.threads = rng_mode, scene_key, threads
    p.exact_marched = 1 if exact else 0
    return p


def adjoint(world, config, dl_dimage, texture=-1, emission_material=-1, threads=0, scene_key=0, exact=True):
    """FP64 path-replay adjoint on the CPU: returns (texel grads (h, w, 3) or
    None, roughness grads per material, emission grad)."""
    p = _params(config, threads=threads, scene_key=scene_key, exact=exact)
    g = np.ascontiguousarray(dl_dimage, dtype=np.float64)
    d = world.desc_struct
    tex = None
    if texture >= 0:
        if texture >= d.n_textures:
            raise ValueError(f"texture {texture} out of range")
        t = d.textures[texture]
        tex = np.zeros((t.height, t.width, 3))
    rough = np.zeros(max(d.n_materials, 1))
    em = C.c_double()
    dp = C.POINTER(C.c_double)
    _check(lib().oracle_adjoint(world.desc, C.byref(p), texture, emission_material, g.ctypes.data_as(dp),
                                tex.ctypes.data_as(dp) if tex is not None else None, rough.ctypes.data_as(dp),
                                C.byref(em)))
    return tex, rough[:d.n_materials], em.value


def _vp(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def aov(world, config, rng_mode=RNG_COUNTER, threads=0, scene_key=0, exact=True, albedo=True):
    """AOV passes on the CPU (render.rs:3508-3584): (depth, normal, albedo);
    depth +inf where the pixel's ray escapes."""
    p = _params(config, rng_mode=rng_mode, threads=threads, scene_key=scene_key, exact=exact)
    fw, fh = max(config.width, 1), max(config.height, 1)
    cw, ch = fw, fh
    if config.crop is not None:
        cw, ch = min(config.crop[2], fw), min(config.crop[3], fh)
    d = np.empty((ch, cw), np.float32)
    n = np.empty((ch, cw, 3), np.uint8)
    a = np.empty((ch, cw, 3), np.uint8) if albedo else None
    _check(lib().oracle_aov(world.desc, C.byref(p), _vp(d), _vp(n), _vp(a)))
    return d, n, a


def post(hdr, exposure=1.0, tone_map=0, auto_exposure=False, fog_depth=None, fog_density=0.0,
         fog_color=(0.8, 0.85, 0.9)):
    """Default post chain on the CPU: (rgba, exposed HDR, exposure used)."""
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    shape = hdr.shape[:-1]
    n = hdr.size // 3
    rgba = np.empty(shape + (4,), np.uint8)
    hx = np.empty(hdr.shape, np.float32)
    fog = None if fog_depth is None else np.ascontiguousarray(fog_depth, dtype=np.float32)
    ex = C.c_double()
    _check(lib().oracle_post(_vp(hdr), n, exposure, tone_map, 1 if auto_exposure else 0, _vp(fog), fog_density,
                             _d3(fog_color), _vp(rgba), _vp(hx), C.byref(ex)))
    return rgba, hx, ex.value


def log_average_luminance(hdr):
    hdr = np.ascontiguousarray(hdr, dtype=np.float32)
    return lib().oracle_log_average_luminance(_vp(hdr), hdr.size // 3)


def tone(mode, x):
    return lib().oracle_tone(mode, x)


def srgb8(x):
    return lib().oracle_srgb8(x)


def rgbe(rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.empty(rgb.shape[:-1] + (4,), np.uint8)
    lib().oracle_rgbe(_vp(rgb), rgb.size // 3, _vp(out))
    return out
