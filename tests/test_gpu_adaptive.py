"""Adaptive sampling and the time limit on the GPU (src/render.rs:1587-1677).

  * plain FP64 kernel vs the oracle (counter RNG): the same samples per pixel
    (equal totals) and relative RMSE below 1e-6 (the FP32 framebuffer's
    rounding, as for the ordinary FP64 render);
  * FP32 wavefront (rounds over the pixels still sampling) vs the plain FP32
    kernel (one pixel per thread, sequential): the same sample values, so
    totals within 0.1 % and images within 1e-6;
  * zero threshold: the wavefront's adaptive image equals its ordinary
    box-filtered render to 1e-6, every sample taken;
  * FP32 wavefront vs FP64 plain kernel: relative RMSE below 1e-3, totals
    within 1 %;
  * time limit: one far away gives the ordinary render bit for bit; one that
    has passed before the first band gives every pixel one centre sample
    (rows = H), matching the oracle's expired render pixel by pixel to 1e-2
    relative on >= 97 % of the pixels (one FP32 path per pixel against an
    FP64 one, no averaging: a path that takes another branch changes its
    whole pixel; another stream or ray would move nearly every pixel by O(1));
    a batch or the plain kernels with a time limit are refused.
"""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import workloads
import oracle

pytestmark = pytest.mark.gpu


def relrmse(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scene(spp=64, w=48, h=40):
    return lm.load_scene(workloads.cornell_box(w, h, spp, 8))


def test_plain_fp64_adaptive_matches_oracle(gpu):
    cfg, cam, world = scene(spp=64, w=24, h=20)
    cfg.adaptive, cfg.adaptive_threshold = True, 0.05
    ref, _, ost = oracle.render(world, cfg)
    g = lm.render(cfg, cam, world, path="plain64")
    print(f"[adaptive] FP64 {g.total_rays} vs oracle {ost['camera_rays']}", flush=True)
    assert g.total_rays == ost["camera_rays"]
    assert g.total_rays < cfg.width * cfg.height * 64
    err = relrmse(g.rows.astype(np.float64), ref)
    print(f"[adaptive] FP64 relRMSE {err:.3e}", flush=True)
    assert err < 1e-6


def test_wavefront_adaptive_equals_plain_fp32(gpu):
    cfg, cam, world = scene(spp=100)
    cfg.adaptive, cfg.adaptive_threshold = True, 0.04
    a = lm.render(cfg, cam, world, path="wavefront")
    b = lm.render(cfg, cam, world, path="plain32")
    print(f"[adaptive] wavefront {a.total_rays} ({a.stats['rounds']} rounds, {a.stats['samples_traced']} traced) "
          f"vs plain FP32 {b.total_rays}", flush=True)
    assert abs(a.total_rays - b.total_rays) <= 1e-3 * b.total_rays
    assert a.stats["samples_traced"] >= a.total_rays
    np.testing.assert_allclose(a.rows, b.rows, rtol=1e-5, atol=1e-6)
    c = lm.render(cfg, cam, world, path="plain64")
    assert abs(a.total_rays - c.total_rays) <= 1e-2 * c.total_rays
    assert relrmse(a.rows.astype(np.float64), c.rows.astype(np.float64)) < 1e-3


def test_zero_threshold_is_the_ordinary_render(gpu):
    cfg, cam, world = scene(spp=36)
    plain = lm.render(cfg, cam, world)
    cfg.adaptive, cfg.adaptive_threshold = True, 0.0
    ad = lm.render(cfg, cam, world)
    assert ad.total_rays == plain.total_rays == cfg.width * cfg.height * 36
    np.testing.assert_allclose(ad.rows, plain.rows, rtol=1e-6, atol=1e-7)


def test_time_limit(gpu):
    cfg, cam, world = scene(spp=16, w=40, h=32)
    base = lm.render(cfg, cam, world)
    cfg.time_limit = 1e9
    far = lm.render(cfg, cam, world)
    assert np.array_equal(far.rows, base.rows) and far.stats["expired_rows"] == 0
    cfg.time_limit = 1e-9
    exp = lm.render(cfg, cam, world)
    print(f"[time limit] expired rows {exp.stats['expired_rows']}, rays {exp.total_rays}", flush=True)
    assert exp.stats["expired_rows"] == cfg.height and exp.total_rays == cfg.width * cfg.height
    ref, _, ost = oracle.render(world, cfg)
    assert ost["camera_rays"] == cfg.width * cfg.height
    err = np.max(np.abs(exp.rows - ref) / np.maximum(np.abs(ref), 1e-3), axis=-1)
    print(f"[time limit] share of pixels within 1e-4 / 1e-3 / 1e-2 of the oracle's: "
          f"{(err <= 1e-4).mean():.4f} / {(err <= 1e-3).mean():.4f} / {(err <= 1e-2).mean():.4f}", flush=True)
    assert (err <= 1e-2).mean() >= 0.97
    # adaptive and expired together: still one centre sample per pixel
    cfg.adaptive = True
    both = lm.render(cfg, cam, world)
    assert both.total_rays == cfg.width * cfg.height
    np.testing.assert_allclose(both.rows, exp.rows, rtol=1e-6, atol=1e-7)
    with pytest.raises(lm.LuminaraError):
        lm.render(cfg, cam, world, path="plain64")
    with pytest.raises(lm.LuminaraError):
        lm.render_batch([cfg, cfg], [world, world])


def test_adaptive_batch(gpu):
    scenes = [scene(spp=64, w=32, h=24), scene(spp=100, w=20, h=20)]
    cfgs = []
    for c, _, _ in scenes:
        c.adaptive, c.adaptive_threshold = True, 0.05
        cfgs.append(c)
    outs, st = lm.render_batch(cfgs, [w for _, _, w in scenes])
    for k, ((c, cam, w), o) in enumerate(zip(scenes, outs)):
        one = lm.render(c, cam, w, scene_key=k)
        np.testing.assert_allclose(o, one.rows, rtol=1e-6, atol=1e-7)
