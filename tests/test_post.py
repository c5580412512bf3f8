"""The oracle's image outputs on the CPU, pinned before the GPU is compared
with them: the default post chain (src/render.rs:2594-2632, :3099-3202,
:3606-3658) against the reference's own known answers for its tone maps,
sRGB quantisation and log-average luminance (src/render.rs:3665-3772) and
against an independent numpy evaluation; the Radiance RGBE pixels
(src/main.rs:964-986) against hand-derived bytes and a decode round trip;
the AOV passes (src/render.rs:3508-3584) against closed-form depth, normal
and albedo bytes of a quad in front of the camera."""
import math

import numpy as np
import pytest

import oracle
import luminara_amd as lm

ACES, REINHARD, FILMIC, NONE = 0, 1, 2, 3


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


# ---- tone curves and sRGB: the reference's known answers --------------------

def test_tone_maps_reference_known_answers():
    t = oracle.tone
    assert t(ACES, 0.0) == 0.0                                    # aces_tonemap_zero
    assert t(REINHARD, 0.0) == 0.0 and abs(t(REINHARD, 1.0) - 0.5) < 1e-6   # reinhard_tonemap_basic
    assert abs(t(FILMIC, 0.0)) < 0.01                             # filmic_tonemap_zero
    # closed form of the ACES fit at 1: (2.51 + 0.03) / (2.43 + 0.59 + 0.14)
    assert t(ACES, 1.0) == pytest.approx(2.54 / 3.16, rel=1e-15)
    for mode in (ACES, FILMIC):                                   # *_tonemap_bounded
        ys = [t(mode, x * 0.5) for x in range(100)]
        assert all(0.0 <= y <= 1.0 for y in ys)
    for mode in (ACES, REINHARD, FILMIC):                         # *_tonemap_monotonic
        ys = [t(mode, x * 0.1) for x in range(1, 100)]
        assert all(b >= a for a, b in zip(ys, ys[1:]))
    ys = [t(REINHARD, float(x)) for x in range(1, 50)]
    assert all(b > a for a, b in zip(ys, ys[1:])) and all(y < 1.0 for y in ys)
    # Hable's curve normalised at its white point 11.2 maps 11.2 to 1
    assert t(FILMIC, 11.2) == pytest.approx(1.0, rel=1e-15)
    assert t(NONE, 0.25) == 0.25 and t(NONE, 7.0) == 1.0 and t(NONE, -3.0) == 0.0


def test_srgb_quantisation_reference_known_answers():
    s = oracle.srgb8
    assert s(0.0) == 0 and s(1.0) == 255 and s(-1.0) == 0         # srgb_black / _white / _clamps_negative
    # 1.055 * 0.5^(1/2.4) - 0.055 = 0.73536; floor(0.73536 * 256) = 188
    assert s(0.5) == 188
    # linear segment: 12.92 * 0.0031308 = 0.04045; floor(0.04045 * 256) = 10
    assert s(0.0031308) == 10 and s(2.0) == 255
    ys = [s(x / 100.0) for x in range(1, 101)]                    # srgb_monotonic
    assert all(b >= a for a, b in zip(ys, ys[1:]))


def test_log_average_luminance_reference_known_answers():
    gray = np.full((10, 10, 3), 0.5, np.float32)                  # log_avg_luminance_basic
    expected = 0.2126 * 0.5 + 0.7152 * 0.5 + 0.0722 * 0.5
    assert abs(oracle.log_average_luminance(gray) - expected) < 0.01
    dark = np.full((10, 10, 3), 0.01, np.float32)                 # log_avg_luminance_dark_scene
    assert oracle.log_average_luminance(dark) < 0.05
    # closed form: geometric mean of (1e-4 + L), negative luminance counts as 0
    img = np.array([[[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [-5.0, 0.0, 0.0]]], np.float32)
    lum = [0.2126, 0.7152 * 2.0, 0.0]
    want = math.exp(sum(math.log(1e-4 + v) for v in lum) / 3)
    assert oracle.log_average_luminance(img) == pytest.approx(want, rel=1e-14)


# ---- the per-pixel chain against an independent numpy evaluation -------------

def np_tone(mode, x):
    with np.errstate(invalid="ignore", over="ignore"):
        if mode == ACES:
            x = np.fmax(x, 0.0)
            y = (x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14)
            return np.where(np.isnan(y), y, np.clip(y, 0.0, 1.0))
        if mode == REINHARD:
            x = np.fmax(x, 0.0)
            return x / (1.0 + x)
        if mode == FILMIC:
            def uc2(v):
                return ((v * (0.15 * v + 0.10 * 0.50) + 0.20 * 0.02) / (v * (0.15 * v + 0.50) + 0.20 * 0.30)) \
                    - 0.02 / 0.30
            x = np.fmax(x, 0.0)
            y = uc2(x) / uc2(11.2)
            return np.where(np.isnan(y), y, np.clip(y, 0.0, 1.0))
        return np.where(np.isnan(x), x, np.clip(x, 0.0, 1.0))


def np_srgb_value(x):
    """s of gamma_correct before quantisation (NaN stays NaN)."""
    with np.errstate(invalid="ignore"):
        x = np.where(np.isnan(x), x, np.clip(x, 0.0, 1.0))
        return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)


def np_bytes(s):
    with np.errstate(invalid="ignore"):
        q = np.where(np.isnan(s), 0.0, np.clip(s, 0.0, 0.999) * 256.0)
    return np.floor(q).astype(np.int64), q


def hdr_sample(rng, n):
    """Linear HDR values over the range a framebuffer holds, plus the edge
    cases of the chain: zeros, negatives, the sRGB knee, NaN, infinities."""
    v = np.concatenate([rng.random(n) * 2.0, np.exp(rng.normal(0.0, 3.0, n)), -rng.random(n // 8),
                        [0.0, 0.0031308, 1.0, 11.2, np.nan, np.inf, -np.inf, 1e30, 1e-30]])
    v = v[: (len(v) // 3) * 3]
    return v.astype(np.float32).reshape(-1, 1, 3)


@pytest.mark.parametrize("mode", [ACES, REINHARD, FILMIC, NONE])
def test_post_chain_matches_numpy(mode):
    rng = np.random.default_rng(7 + mode)
    hdr = hdr_sample(rng, 30000)
    exposure = 1.7
    rgba, hx, used = oracle.post(hdr, exposure=exposure, tone_map=mode)
    assert used == exposure and np.all(rgba[..., 3] == 255)
    c = hdr.astype(np.float64) * exposure
    assert np.array_equal(hx, c.astype(np.float32), equal_nan=True)
    want, q = np_bytes(np_srgb_value(np_tone(mode, c)))
    got = rgba[..., :3].astype(np.int64)
    # numpy's power and the C library's may round the last bit apart: allowed
    # only where the scaled value sits on a quantisation step
    near = (np.abs(q - np.round(q)) < 1e-9) & (q > 0.5)
    assert np.all((got == want) | near)
    assert near.mean() < 1e-3
    # Rust semantics at the edges: an infinite channel is NaN through ACES and
    # Hable's curve, and NaN quantises to 0 (`as u8`)
    inf_idx = np.where(np.isposinf(hdr.reshape(-1)))[0][0]
    if mode in (ACES, FILMIC, REINHARD):
        assert rgba.reshape(-1, 4)[inf_idx // 3, inf_idx % 3] == 0


def test_post_auto_exposure_and_fog():
    rng = np.random.default_rng(11)
    hdr = (rng.random((20, 30, 3)) * 3.0).astype(np.float32)
    _, _, used = oracle.post(hdr, exposure=0.5, auto_exposure=True)
    assert used == pytest.approx(0.5 * 0.18 / oracle.log_average_luminance(hdr), rel=1e-15)
    # a black frame: the average is 1e-4 (> 1e-6), so the key factor is 1800
    _, _, used = oracle.post(np.zeros((4, 4, 3), np.float32), exposure=1.0, auto_exposure=True)
    assert used == pytest.approx(0.18 / 1e-4, rel=1e-12)
    # fog: escaped rays take the fog colour, depth 0 leaves the pixel alone
    depth = np.where(rng.random((20, 30)) < 0.3, np.inf, rng.random((20, 30)) * 20.0).astype(np.float32)
    depth[0, 0] = 0.0
    fc = (0.8, 0.85, 0.9)
    rgba, _, _ = oracle.post(hdr, exposure=1.0, tone_map=NONE, fog_depth=depth, fog_density=0.1, fog_color=fc)
    f = 1.0 - np.exp(-0.1 * depth.astype(np.float64))
    c = hdr.astype(np.float64) * (1.0 - f[..., None]) + np.array(fc) * f[..., None]
    c = np.where(np.isinf(depth)[..., None], np.array(fc), c)
    want, q = np_bytes(np_srgb_value(np_tone(NONE, c)))
    near = (np.abs(q - np.round(q)) < 1e-9) & (q > 0.5)
    assert np.all((rgba[..., :3] == want) | near)
    plain, _, _ = oracle.post(hdr[:1, :1], exposure=1.0, tone_map=NONE)
    assert np.array_equal(rgba[0, 0], plain[0, 0])


# ---- Radiance RGBE pixels ----------------------------------------------------

def test_rgbe_known_answers():
    px = np.array([[1.0, 1.0, 1.0], [0.5, 0.25, 0.0], [3.0, 1.0, 0.0], [0.0, 0.0, 0.0], [-1.0, 2.0, 0.0],
                   [1e-33, 0.0, 0.0], [np.nan, 1.0, 0.0], [np.inf, 0.0, 0.0]], np.float32)
    got = oracle.rgbe(px)
    # max 1: e = ceil(log2 1) = 0, 1 * 256 clamps to 255, exponent byte 128
    assert list(got[0]) == [255, 255, 255, 128]
    # max 0.5: e = -1, scale 512: 256 -> 255, 128, 0; exponent byte 127
    assert list(got[1]) == [255, 128, 0, 127]
    # max 3: e = 2, scale 64: 192, 64, 0; exponent byte 130
    assert list(got[2]) == [192, 64, 0, 130]
    assert list(got[3]) == [0, 0, 0, 0] and list(got[5]) == [0, 0, 0, 0]
    # negative channels clamp to 0 (f32::max), max 2: e = 1, scale 128
    assert list(got[4]) == [0, 255, 0, 129]
    # NaN loses against 0 in f32::max
    assert list(got[6]) == [0, 255, 0, 128]
    # inf: e saturates to i32::MAX, the scale is 0, inf * 0 = NaN and
    # f32::min(NaN, 255) = 255; (i32::MAX + 128) wraps to 0x8000007F -> 127
    assert list(got[7]) == [255, 0, 0, 127]


def test_rgbe_round_trip_and_powers_of_two():
    rng = np.random.default_rng(3)
    v = np.exp(rng.normal(0.0, 4.0, (5000, 3))).astype(np.float32)
    e = oracle.rgbe(v).astype(np.float64)
    dec = e[:, :3] * np.ldexp(1.0, (e[:, 3:] - 128).astype(int)) / 256.0
    m = v.max(axis=1, keepdims=True).astype(np.float64)
    assert np.all(np.abs(dec - v) <= m / 128.0)
    # the shared exponent comes from the f32 log2: at and next to powers of
    # two it must equal ceil of the correctly rounded f32 log2 (what the GPU
    # computes, from an FP64 log2 rounded once)
    k = np.arange(-100, 120)
    p2 = np.ldexp(np.float32(1.0), k).astype(np.float32)
    xs = np.concatenate([p2, np.nextafter(p2, np.float32(np.inf)), np.nextafter(p2, np.float32(0)),
                         p2 * np.float32(1.0000005)]).astype(np.float32)
    xs = xs[xs > 1e-32]
    got = oracle.rgbe(np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], axis=1))
    want = np.ceil(np.log2(xs.astype(np.float64)).astype(np.float32)).astype(np.int64) + 128
    assert np.array_equal(got[:, 3].astype(np.int64), want & 0xFF)


# ---- AOV passes: closed forms on a quad facing the camera --------------------

def quad_scene(material, width=32, height=24, aperture=0.0):
    toml = f"""
[render]
width = {width}
height = {height}
samples = 1
max_depth = 4
seed = 42
background = {{ type = "black" }}

[camera]
look_from = [0.0, 0.0, -10.0]
look_at = [0.0, 0.0, 0.0]
vfov = 40.0
aperture = {aperture}
focus_dist = 1.0

[[quad]]
q = [-2.0, -1.5, 5.0]
u = [4.0, 0.0, 0.0]
v = [0.0, 3.0, 0.0]
material = {material}
"""
    return lm.load_scene(toml)


def closed_form_depth(cam, width, height):
    o = np.array(cam.origin[:])
    ll, hz, vt = (np.array(a[:]) for a in (cam.lower_left, cam.horizontal, cam.vertical))
    j, i = np.mgrid[0:height, 0:width]
    u = (i + 0.5) / (width - 1)
    v = ((height - 1 - j) + 0.5) / (height - 1)
    d = ll + u[..., None] * hz + v[..., None] * vt - o
    t = (5.0 - o[2]) / d[..., 2]
    p = o + t[..., None] * d
    a, b = (p[..., 0] + 2.0) / 4.0, (p[..., 1] + 1.5) / 3.0
    inside = (a >= 0) & (a <= 1) & (b >= 0) & (b <= 1)
    edge = (np.minimum(np.abs(a), np.abs(a - 1)) < 1e-9) | (np.minimum(np.abs(b), np.abs(b - 1)) < 1e-9)
    return t, inside, edge


def test_aov_closed_forms_lambertian():
    cfg, cam, world = quad_scene('{ type = "lambertian", color = [0.5, 0.25, 1.0] }')
    depth, normal, albedo = oracle.aov(world, cfg)
    t, inside, edge = closed_form_depth(cam, cfg.width, cfg.height)
    assert inside.any() and (~inside).any()
    ok = ~edge
    assert np.array_equal(np.isfinite(depth)[ok], inside[ok])
    assert np.allclose(depth[inside & ok], t[inside & ok].astype(np.float32), rtol=1e-6, atol=0)
    # the facing normal is (0, 0, -1): bytes floor(127.5), floor(127.5), 0
    assert np.all(normal[inside & ok] == [127, 127, 0]) and np.all(normal[~inside & ok] == 0)
    # Lambertian attenuation = albedo: floor(0.5*255), floor(0.25*255), 255
    assert np.all(albedo[inside & ok] == [127, 63, 255]) and np.all(albedo[~inside & ok] == 0)
    # a zero-aperture camera and a Lambertian albedo use no random number that
    # changes a byte: the reference's per-pixel xoshiro stream gives the same
    d2, n2, a2 = oracle.aov(world, cfg, rng_mode=oracle.RNG_REFERENCE)
    assert np.array_equal(depth, d2) and np.array_equal(normal, n2) and np.array_equal(albedo, a2)
    # a crop is the same block of the full frame
    cfg.crop = (5, 3, 17, 11)
    dc, nc, ac = oracle.aov(world, cfg)
    assert np.array_equal(dc, depth[3:14, 5:22]) and np.array_equal(ac, albedo[3:14, 5:22])


@pytest.mark.parametrize("mat,want", [
    ('{ type = "emissive", color = [0.5, 0.2, 0.0], intensity = 1.0 }', [127, 51, 0]),
    ('{ type = "emissive", color = [3.0, 1.5, 0.3], intensity = 10.0 }',
     [255, int(15.0 / 30.0 * 255.0), int(3.0 / 30.0 * 255.0)]),
    ('{ type = "metal", color = [0.9, 0.6, 0.3], fuzz = 0.0 }',
     [int(0.9 * 255.0), int(0.6 * 255.0), int(0.3 * 255.0)]),
    ('{ type = "dielectric", refraction_index = 1.5 }', [255, 255, 255]),
])
def test_aov_albedo_closed_forms(mat, want):
    cfg, cam, world = quad_scene(mat)
    _, _, albedo = oracle.aov(world, cfg)
    _, inside, edge = closed_form_depth(cam, cfg.width, cfg.height)
    assert np.all(albedo[inside & ~edge] == want)
