"""The CPU FP64 path-replay adjoint (oracle) against central finite
differences with common random numbers (CPU only).

The adjoint differentiates the expected image (sampled directions held
fixed, Russian roulette through its survival probability), while a finite
difference with common random numbers differentiates the sampled image, so
they agree exactly only where the sampled image is smooth in the parameter:
  * emission: the image is linear in the intensity -- exact;
  * texels with no Russian roulette on the path (depth >= 47 remaining):
    the image is linear in the albedo -- exact;
and otherwise in expectation, checked over several seeds within the
standard errors:
  * GGX roughness (the directions GGX sampling draws move with roughness);
  * texels with roulette on every bounce (its decisions move with albedo).
A linear functional sum(G * I) with fixed weights G stands in for the loss,
so the finite difference has no bias of its own.
"""
import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import workloads
import oracle

H_EXACT = 2.0 ** -10  # exact in FP32 texels


def functional(world, cfg, G):
    I, _, _ = oracle.render(world, cfg, threads=8, want_var=False)
    return float(np.sum(G * I))


def test_emission_gradient_is_exact():
    def scene(intensity):
        toml, setup = workloads.inverse_rendering_scene(24, 24, 16, 6, 42, tex_res=8, n_spheres=4,
                                                        intensity=intensity)
        cfg, _, w = lm.load_scene(toml)
        return cfg, w, setup(w)
    cfg, w, info = scene(1.0)
    G = np.random.default_rng(1).normal(size=(24, 24, 3))
    _, _, ge = oracle.adjoint(w, cfg, G, emission_material=info["emission_material"], threads=8)
    h = 1e-3
    fp = functional(*scene(1.0 + h)[1::-1], G)
    fm = functional(*scene(1.0 - h)[1::-1], G)
    fd = (fp - fm) / (2 * h)
    assert abs(ge) > 1.0
    assert ge == pytest.approx(fd, rel=1e-7)


NO_ROULETTE = """
[render]
width = 24
height = 24
samples = 16
max_depth = 50
[camera]
look_from = [0.0, 3.0, 0.01]
look_at = [0.0, 0.0, 0.0]
vfov = 60.0
[[quad]]
q = [-2.0, 0.0, -2.0]
u = [4.0, 0.0, 0.0]
v = [0.0, 0.0, 4.0]
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
[[quad]]
q = [-0.3, 2.0, -0.3]
u = [0.6, 0.0, 0.0]
v = [0.0, 0.0, 0.6]
material = { type = "emissive", color = [1.0, 0.9, 0.8], intensity = 5.0 }
"""


def test_texel_gradient_is_exact_without_roulette():
    # a textured floor under the sky and a small NEE light: every path is
    # camera -> floor -> sky or light, at remaining depth 50 and 49, where the
    # reference applies no roulette; the image is linear in every texel
    cfg, _, w = lm.load_scene(NO_ROULETTE)
    rng = np.random.default_rng(3)
    tex0 = (rng.integers(0, 256, size=(4, 4, 3)) / 255.0).astype(np.float32)
    t = w.add_texture(tex0)
    m = w.materials()[0]
    m.texture = t
    w.set_material(0, m)
    G = rng.normal(size=(24, 24, 3))
    gt, _, ge = oracle.adjoint(w, cfg, G, texture=t, emission_material=1, threads=8)
    for (j, i, c) in [(0, 0, 0), (1, 2, 1), (3, 3, 2), (2, 1, 0)]:
        vals = []
        for s in (1, -1):
            tx = tex0.copy()
            tx[j, i, c] += np.float32(s * H_EXACT)
            w.set_texture(t, tx)
            vals.append(functional(w, cfg, G))
        w.set_texture(t, tex0)
        fd = (vals[0] - vals[1]) / (2 * H_EXACT)
        assert gt[j, i, c] == pytest.approx(fd, rel=1e-8, abs=1e-9), (j, i, c)
    assert ge != 0.0


def _seeds_agree(fd, adj, rel):
    fd, adj = np.asarray(fd), np.asarray(adj)
    se = np.sqrt(fd.var(ddof=1) / len(fd) + adj.var(ddof=1) / len(adj))
    diff = abs(fd.mean() - adj.mean())
    assert diff <= 3.0 * se + 1e-12, (fd.mean(), adj.mean(), se)
    assert diff <= rel * abs(fd.mean()), (fd.mean(), adj.mean())


def test_roughness_gradient_in_expectation():
    # one GGX bounce under a gradient sky (no roulette at remaining depth 50,
    # a smooth background): the central difference with common random
    # numbers is resolved to 0.3 %, and the adjoint agrees within 1 %
    toml = """
[render]
width = 16
height = 16
samples = 16384
max_depth = 50
background = { type = "gradient", bottom = [1.0, 0.2, 0.1], top = [0.1, 0.3, 1.0] }
[camera]
look_from = [0.0, 3.0, 2.0]
look_at = [0.0, 0.0, 0.0]
vfov = 40.0
[[quad]]
q = [-3.0, 0.0, -3.0]
u = [6.0, 0.0, 0.0]
v = [0.0, 0.0, 6.0]
material = { type = "pbr", color = [0.8, 0.6, 0.4], roughness = 0.35, metallic = 0.0 }
"""
    cfg, _, w = lm.load_scene(toml)
    G = np.random.default_rng(3).normal(size=(16, 16, 3)) + 1.0
    fds, adjs = [], []
    for seed in range(1, 17):
        cfg.seed = seed
        _, gr, _ = oracle.adjoint(w, cfg, G, threads=8)
        vals = []
        for s in (1, -1):
            m = w.materials()[0]
            m.roughness = 0.35 + s * 1e-3
            w.set_material(0, m)
            vals.append(functional(w, cfg, G))
        m = w.materials()[0]
        m.roughness = 0.35
        w.set_material(0, m)
        fds.append((vals[0] - vals[1]) / 2e-3)
        adjs.append(gr[0])
    fd = np.asarray(fds)
    assert fd.std(ddof=1) / np.sqrt(len(fd)) <= 3e-3 * abs(fd.mean())
    _seeds_agree(fds, adjs, rel=0.01)


def test_texel_gradient_with_roulette_in_expectation():
    # between two Lambertian planes with a ceiling light and max_depth 5:
    # roulette on every bounce, its survival probability max(albedo) moves
    # with the texels (the NEE term is not divided by it)
    toml = """
[render]
width = 16
height = 16
samples = 4096
max_depth = 5
background = { type = "solid", color = [0.3, 0.4, 0.5] }
[camera]
look_from = [0.0, 1.0, 4.0]
look_at = [0.0, 1.0, 0.0]
vfov = 50.0
[[quad]]
q = [-3.0, 0.0, -3.0]
u = [6.0, 0.0, 0.0]
v = [0.0, 0.0, 6.0]
material = { type = "lambertian", color = [0.5, 0.5, 0.5] }
[[quad]]
q = [-3.0, 2.0, -3.0]
u = [6.0, 0.0, 0.0]
v = [0.0, 0.0, 6.0]
material = { type = "lambertian", color = [0.6, 0.7, 0.4] }
[[quad]]
q = [-0.5, 1.95, -0.5]
u = [1.0, 0.0, 0.0]
v = [0.0, 0.0, 1.0]
material = { type = "emissive", color = [1.0, 1.0, 1.0], intensity = 6.0 }
"""
    cfg, _, w = lm.load_scene(toml)
    tex0 = np.full((4, 4, 3), np.float32(0.5), dtype=np.float32)
    tex0[..., 0] = np.float32(0.625)
    t = w.add_texture(tex0)
    m = w.materials()[0]
    m.texture = t
    w.set_material(0, m)
    G = np.random.default_rng(3).normal(size=(16, 16, 3)) + 1.0
    fds, adjs = [], []
    for seed in range(1, 9):
        cfg.seed = seed
        gt, _, _ = oracle.adjoint(w, cfg, G, texture=t, threads=8)
        vals = []
        for s in (1, -1):
            w.set_texture(t, tex0 + np.float32(s * H_EXACT))
            vals.append(functional(w, cfg, G))
        w.set_texture(t, tex0)
        fds.append((vals[0] - vals[1]) / (2 * H_EXACT))
        adjs.append(float(gt.sum()))
    _seeds_agree(fds, adjs, rel=0.02)


def test_adjoint_rejects_bad_parameters():
    cfg, _, w = lm.load_scene(NO_ROULETTE)
    G = np.zeros((24, 24, 3))
    with pytest.raises(ValueError):
        oracle.adjoint(w, cfg, G, texture=5)
