"""Image outputs on the GPU against the CPU oracle: the default post chain
(exposure, automatic exposure, depth fog, tone curve, sRGB, 8-bit), the
Radiance RGBE pixels and the AOV passes, through the C ABI.

Tolerances:
  * post chain: bytes equal (the sRGB step is read off thresholds found
    with the host's own pow); with depth fog, equal except where the value
    before truncation lies within 1e-9 of a quantisation step (the GPU's and
    the host's exp may round the last bit apart there; at most 1e-3 of the
    channels); exposed HDR bit-exact; automatic exposure within 1e-12
    relative (the GPU sums the logarithms in another order).
  * RGBE: bit-exact.
  * AOV, FP64 trace: depth within 1e-9 relative where both hit; hit masks,
    normal and albedo bytes equal on >= 99.9 % of the pixels (grazing
    edges and quantisation steps), not counting exact ties: pixels whose
    ray meets two primitives at the same t (a shared mesh edge on the
    camera's symmetry plane) take the later primitive in the reference and
    the first here -- the depth agrees to 1e-12 and the bytes are another
    primitive's (at most 1 % of the pixels).
  * AOV, FP32 trace: depth within max(1e-5, 2 eps32 (|o| + |ll| + |h| + |v|) / |d|)
    + 8 eps32 M / (t |d.n|) relative -- the camera's FP32 origin and corner
    cancel in d = ll + s h + t v - o, so a camera far from the origin cannot
    place d more precisely, and FP32 cannot place a hit point more precisely
    than eps32 M (M the largest coordinate of origin and hit point), which
    moves t by that over |d.n| (n the unit normal of the oracle's hit) -- on
    >= 99.5 % of the pixels both hit, normal bytes within 1 on >= 99 % of
    them and albedo bytes equal on >= 99 % of all pixels.
  * lmn_render_image: bit-exact against lmn_render, lmn_aov, lmn_post and
    lmn_encode_rgbe called one by one.
"""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import workloads
import oracle
from hipmem import DeviceBuffer

pytestmark = pytest.mark.gpu
ACES, REINHARD, FILMIC, NONE = 0, 1, 2, 3
EPS32 = 2.0 ** -24


def step_distance(hdr, exposure, tone, fog_depth=None, density=0.0, fc=(0.8, 0.85, 0.9)):
    """Distance of each channel's scaled sRGB value from the nearest integer
    (numpy evaluation of the chain, only used to find values on a step)."""
    c = hdr.astype(np.float64) * exposure
    with np.errstate(invalid="ignore", over="ignore"):
        if fog_depth is not None:
            f = 1.0 - np.exp(-density * fog_depth.astype(np.float64))[..., None]
            c = c * (1.0 - f) + np.array(fc) * f
        x = np.fmax(c, 0.0) if tone != NONE else c
        if tone == ACES:
            x = (x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14)
        elif tone == REINHARD:
            x = x / (1.0 + x)
        elif tone == FILMIC:
            def uc2(v):
                return ((v * (0.15 * v + 0.05) + 0.004) / (v * (0.15 * v + 0.5) + 0.06)) - 0.02 / 0.3
            x = uc2(x) / uc2(11.2)
        x = np.clip(x, 0.0, 1.0)
        s = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)
        q = np.clip(s, 0.0, 0.999) * 256.0
    d = np.abs(q - np.round(q))
    return np.where(np.isnan(d) | (q < 0.5), 1.0, d)


def compare_bytes(got, want, dist):
    diff = got[..., :3] != want[..., :3]
    assert np.all(got[..., 3] == 255)
    assert np.all(~diff | (dist < 1e-9)), f"{int(diff.sum())} channels differ off a quantisation step"
    assert diff.mean() <= 1e-3


def hdr_frame(rng, h, w):
    v = np.exp(rng.normal(-0.5, 2.5, (h, w, 3)))
    v[rng.random((h, w)) < 0.02] = 0.0
    flat = v.reshape(-1)
    flat[:8] = [np.nan, np.inf, -1.0, 1e30, 1e-30, 0.0031308, 11.2, 1.0]
    return v.astype(np.float32)


@pytest.mark.parametrize("tone", [ACES, REINHARD, FILMIC, NONE])
def test_post_matches_oracle(gpu, tone):
    rng = np.random.default_rng(100 + tone)
    hdr = hdr_frame(rng, 181, 243)
    cfg = lm.RenderConfig(exposure=1.3, tone_map=tone)
    rgba, hx, used = lm.post(hdr, cfg, want_hdr=True)
    o_rgba, o_hx, o_used = oracle.post(hdr, exposure=1.3, tone_map=tone)
    print(f"[post] tone {tone} {hdr.shape}", flush=True)
    assert used == o_used == 1.3
    assert np.array_equal(hx, o_hx, equal_nan=True)
    assert np.array_equal(rgba, o_rgba)
    # the per-pixel path (device buffers not 16-B aligned, a pixel count
    # that is not a multiple of 4) gives the same bytes
    n = hdr.shape[0] * hdr.shape[1]
    d_hdr = DeviceBuffer.from_array(hdr.reshape(-1))
    d_rgba, d_hx = DeviceBuffer(n * 4), DeviceBuffer(n * 12)
    lm.post_device(d_hdr.addr(12), n - 1, 1, cfg, 0, d_rgba.addr(4), d_hx.addr(12))
    assert np.array_equal(d_rgba.to_array(np.uint8, (n, 4))[1:], rgba.reshape(-1, 4)[1:])
    assert np.array_equal(d_hx.to_array(np.float32, (n, 3))[1:], hx.reshape(-1, 3)[1:], equal_nan=True)


def test_post_auto_exposure_and_fog_match_oracle(gpu):
    rng = np.random.default_rng(5)
    hdr = hdr_frame(rng, 270, 480)
    hdr[~np.isfinite(hdr)] = 0.5   # a NaN or infinite pixel would make the log-average NaN / inf
    depth = np.where(rng.random((270, 480)) < 0.2, np.inf, rng.random((270, 480)) * 30.0).astype(np.float32)
    cfg = lm.RenderConfig(exposure=0.8, tone_map=ACES, auto_exposure=True, depth_fog=0.07,
                          depth_fog_color=(0.7, 0.8, 0.95))
    rgba, hx, used = lm.post(hdr, cfg, fog_depth=depth, want_hdr=True)
    o_rgba, o_hx, o_used = oracle.post(hdr, exposure=0.8, tone_map=ACES, auto_exposure=True, fog_depth=depth,
                                       fog_density=0.07, fog_color=(0.7, 0.8, 0.95))
    print(f"[post] auto exposure {used} vs {o_used}", flush=True)
    assert used == pytest.approx(o_used, rel=1e-12)
    # the exposure itself may differ in its last bits: compare through the
    # oracle run at the GPU's exposure
    o_rgba, o_hx, _ = oracle.post(hdr, exposure=used, tone_map=ACES, fog_depth=depth, fog_density=0.07,
                                  fog_color=(0.7, 0.8, 0.95))
    assert np.array_equal(hx, o_hx)
    compare_bytes(rgba, o_rgba, step_distance(hdr, used, ACES, depth, 0.07, (0.7, 0.8, 0.95)))


def test_rgbe_matches_oracle(gpu):
    rng = np.random.default_rng(9)
    v = np.exp(rng.normal(0.0, 6.0, (100000, 3))).astype(np.float32)
    v[rng.random(100000) < 0.05] *= -1.0
    k = np.arange(-100, 127)
    p2 = np.ldexp(np.float32(1.0), k).astype(np.float32)
    edge = np.concatenate([p2, np.nextafter(p2, np.float32(np.inf)), np.nextafter(p2, np.float32(0))])
    extra = np.stack([edge, edge * np.float32(0.5), np.zeros_like(edge)], axis=1).astype(np.float32)
    special = np.array([[0, 0, 0], [np.nan, 1, 2], [np.inf, 0, 0], [1e-33, 0, 0], [3e38, 1, 0]], np.float32)
    data = np.concatenate([v, extra, special]).astype(np.float32)
    print(f"[rgbe] {len(data)} pixels", flush=True)
    want = oracle.rgbe(data)
    assert np.array_equal(lm.encode_rgbe(data), want)
    d_in, d_out = DeviceBuffer.from_array(data.reshape(-1)), DeviceBuffer(len(data) * 4)
    lm.encode_rgbe_device(d_in.addr(12), len(data) - 1, d_out.addr(4))  # per-pixel path
    assert np.array_equal(d_out.to_array(np.uint8, (len(data), 4))[1:], want[1:])


def ray_length(cam, cfg):
    """|d| of each pixel-centre camera ray (zero aperture part)."""
    o, ll, hz, vt = (np.array(a[:]) for a in (cam.origin, cam.lower_left, cam.horizontal, cam.vertical))
    j, i = np.mgrid[0:cfg.height, 0:cfg.width]
    u = (i + 0.5) / (cfg.width - 1)
    v = ((cfg.height - 1 - j) + 0.5) / (cfg.height - 1)
    return np.linalg.norm(ll + u[..., None] * hz + v[..., None] * vt - o, axis=-1)


def hit_conditioning(world, cam, cfg):
    """8 eps32 M / (t |d.n|) per pixel from the oracle's hit of the pinhole
    pixel-centre ray (0 where it escapes)."""
    o, ll, hz, vt = (np.array(a[:]) for a in (cam.origin, cam.lower_left, cam.horizontal, cam.vertical))
    j, i = np.mgrid[0:cfg.height, 0:cfg.width]
    u = (i + 0.5) / (cfg.width - 1)
    v = ((cfg.height - 1 - j) + 0.5) / (cfg.height - 1)
    d = (ll + u[..., None] * hz + v[..., None] * vt - o).reshape(-1, 3)
    rays = np.concatenate([np.broadcast_to(o, d.shape), d, np.full((len(d), 1), 0.001),
                           np.full((len(d), 1), np.inf)], axis=1)
    h = oracle.trace(world, rays)
    t = h["t"]
    p = o + d * t[:, None]
    M = np.maximum(np.abs(o).max(), np.abs(p).max(axis=1))
    dn = np.maximum(np.abs(np.einsum("ij,ij->i", d, h["normal"])), 1e-30)
    out = np.where(h["prim"] >= 0, 8 * EPS32 * M / (np.maximum(t, 1e-30) * dn), 0.0)
    return out.reshape(cfg.height, cfg.width)


def camera_magnitude(cam):
    return sum(np.abs(np.array(a[:])).max() for a in (cam.origin, cam.lower_left, cam.horizontal, cam.vertical))


def aov_scenes():
    toml, add_mesh = workloads.mesh_heavy(96, 54, 1, 4, freq=16, n_tori=64, n_superellipsoids=16)
    return {
        "cornell": (workloads.cornell_box(64, 64, 1, 8), None),
        "spheres": (workloads.random_spheres(80, 45, 1, 8, grid=11), None),
        "mesh": (toml, add_mesh),
    }


@pytest.mark.parametrize("name", ["cornell", "spheres", "mesh"])
def test_aov_matches_oracle(gpu, name):
    toml, add = aov_scenes()[name]
    cfg, cam, world = lm.load_scene(toml)
    if add is not None:
        add(world)
    od, on, oa = oracle.aov(world, cfg)
    hit_o = np.isfinite(od)
    for path, fp32 in (("plain64", False), ("wavefront", True)):
        print(f"[aov] {name} {path} {cfg.width}x{cfg.height}", flush=True)
        d, n, a = lm.aov(cfg, world, path=path)
        hit_g = np.isfinite(d)
        both = hit_o & hit_g
        assert (hit_o != hit_g).mean() <= (5e-3 if fp32 else 1e-3)
        with np.errstate(invalid="ignore"):
            rel_all = np.where(both, np.abs(d.astype(np.float64) - od) / od, np.inf)
        rel = rel_all[both]
        if fp32:
            tol = (np.maximum(1e-5, 2 * EPS32 * camera_magnitude(cam) / ray_length(cam, cfg))
                   + hit_conditioning(world, cam, cfg))[both]
            print(f"[aov] {name} FP32: {(rel <= tol).mean():.5f} of depths within tolerance", flush=True)
            assert (rel <= tol).mean() >= 0.995
            assert (np.abs(n[both].astype(int) - on[both].astype(int)).max(axis=-1) <= 1).mean() >= 0.99
            assert (a == oa).all(axis=-1).mean() >= 0.99
        else:
            assert (rel <= 1e-9).mean() >= 0.999
            differ = ~(n == on).all(axis=-1) | ~(a == oa).all(axis=-1)
            tie = rel_all <= 1e-12
            print(f"[aov] {name} FP64: {int(differ.sum())} pixels differ, {int((differ & tie).sum())} of them ties",
                  flush=True)
            assert (differ & ~tie).mean() <= 1e-3
            assert (differ & tie).mean() <= 1e-2
        assert np.all(n[~hit_g] == 0) and np.all(a[~hit_g] == 0)


def test_render_image_equals_its_parts(gpu):
    cfg, cam, world = lm.load_scene(workloads.cornell_box(72, 56, 4, 6))
    cfg.auto_exposure, cfg.depth_fog, cfg.tone_map = True, 0.002, FILMIC
    img = lm.render_image(cfg, world, save_hdr=True, save_depth=True, save_normals=True, save_albedo=True)
    print(f"[image] post {img.post_ms:.3f} ms exposure {img.exposure}", flush=True)
    rows = lm.render(cfg, cam, world).rows
    d, n, a = lm.aov(cfg, world)
    rgba, hx, used = lm.post(rows, cfg, fog_depth=d, want_hdr=True)
    assert img.exposure == used
    assert np.array_equal(img.pixels, rgba) and np.array_equal(img.hdr_data, hx)
    assert np.array_equal(img.rgbe, lm.encode_rgbe(hx))
    assert np.array_equal(img.depth_pass, np.where(np.isinf(d), 0.0, d).astype(np.float32))
    assert np.array_equal(img.normal_pass, n) and np.array_equal(img.albedo_pass, a)
    # only the RGBA image asked for: the same pixels
    cfg.depth_fog = 0.0
    img2 = lm.render_image(cfg, world)
    rgba2, _, _ = lm.post(rows, cfg)
    assert np.array_equal(img2.pixels, rgba2) and img2.hdr_data is None
