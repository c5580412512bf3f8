"""The oracle's adaptive and time-limited pixel loops (src/render.rs:1587-1677)
on the CPU, pinned by what the stopping rule fixes: a zero threshold never
stops (the unweighted mean of every sample, which is the box-filtered render),
a huge one stops every pixel at max(spp/4, 16) samples, a frame of one
constant colour stops there with that colour, the ray count falls as the
threshold rises, and a time limit that has passed before the first row
renders every pixel with one sample through its centre (closed form on a
sky-only scene)."""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import workloads
import oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def cornell(spp=64, w=12, h=10):
    cfg, cam, world = lm.load_scene(workloads.cornell_box(w, h, spp, 6))
    return cfg, cam, world


def test_zero_threshold_is_the_box_filtered_render():
    cfg, cam, world = cornell(spp=36)
    ref, _, st_ref = oracle.render(world, cfg)
    cfg.adaptive, cfg.adaptive_threshold = True, 0.0
    ad, _, st = oracle.render(world, cfg)
    assert st["camera_rays"] == st_ref["camera_rays"] == cfg.width * cfg.height * 36
    np.testing.assert_allclose(ad, ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("spp,per_pixel", [(64, 16), (256, 64), (4, 4), (17, 16)])
def test_huge_threshold_stops_at_the_minimum(spp, per_pixel):
    # min = max(spp_a / 4, 16); with spp_a < 16 the test is never reached
    # (spp 17 rounds up to 25 = 5 x 5 strata: min 16 <= 25)
    cfg, cam, world = cornell(spp=spp, w=6, h=5)
    cfg.adaptive, cfg.adaptive_threshold = True, 1e30
    _, _, st = oracle.render(world, cfg)
    assert st["camera_rays"] == cfg.width * cfg.height * per_pixel


def test_rays_fall_as_the_threshold_rises():
    cfg, cam, world = cornell(spp=100)
    cfg.adaptive = True
    rays = []
    for thr in (0.0, 0.01, 0.05, 0.2, 1e30):
        cfg.adaptive_threshold = thr
        rays.append(oracle.render(world, cfg)[2]["camera_rays"])
    assert all(a >= b for a, b in zip(rays, rays[1:]))
    assert rays[0] == cfg.width * cfg.height * 100 and rays[-1] == cfg.width * cfg.height * 25
    assert rays[1] > rays[-1]


SKY = """
[render]
width = 9
height = 7
samples = 64
max_depth = 4
seed = 5
background = { type = "solid", color = [0.25, 0.5, 0.75] }

[camera]
look_from = [0.0, 0.0, 0.0]
look_at = [0.0, 0.0, -1.0]
vfov = 60.0
aperture = 0.0
focus_dist = 1.0
"""


def test_constant_frame_stops_at_the_minimum_with_its_colour():
    cfg, cam, world = lm.load_scene(SKY)
    cfg.adaptive = True
    mean, _, st = oracle.render(world, cfg)
    assert st["camera_rays"] == 9 * 7 * 16      # zero variance: the first test stops it
    assert np.all(mean == np.array([0.25, 0.5, 0.75]))


def test_expired_time_limit_takes_one_centre_sample():
    toml = SKY.replace('background = { type = "solid", color = [0.25, 0.5, 0.75] }', "background = { type = \"sky\" }")
    cfg, cam, world = lm.load_scene(toml)
    cfg.time_limit = 1e-12
    mean, _, st = oracle.render(world, cfg)
    assert st["camera_rays"] == 9 * 7
    # closed form: the sky gradient of the unit direction through each pixel centre
    o, ll, hz, vt = (np.array(a[:]) for a in (cam.origin, cam.lower_left, cam.horizontal, cam.vertical))
    j, i = np.mgrid[0:7, 0:9]
    u = (i + 0.5) / 8
    v = ((6 - j) + 0.5) / 6
    d = ll + u[..., None] * hz + v[..., None] * vt - o
    t = 0.5 * (d[..., 1] / np.linalg.norm(d, axis=-1) + 1.0)
    sky = (1.0 - t)[..., None] * np.ones(3) + t[..., None] * np.array([0.5, 0.7, 1.0])
    np.testing.assert_allclose(mean, sky, rtol=1e-12)
    # no limit (or one far away) is the ordinary render
    cfg.time_limit = 0.0
    a = oracle.render(world, cfg)[0]
    cfg.time_limit = 1e9
    b = oracle.render(world, cfg)[0]
    assert np.array_equal(a, b)
