"""Golden ray-batch fixtures (tests/golden/rays_*.npz, written by
tests/golden/make_golden.py from the CPU FP64 oracle).

CPU: the oracle still reproduces every fixture (a regression guard on the
checker itself).  GPU: the FP64 kernel matches the fixtures to 1e-9 relative
in t with identical prim ids, the FP32 kernel to the FP32 conditioning bound,
in both intersect modes (the reference's marchers on rays exact in FP32); no
oracle is needed at run time.
"""
import glob
import os

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "rays_*.npz")))
MODES = {"reference": _abi.ISECT_REFERENCE, "exact": _abi.ISECT_EXACT}


def load(path):
    z = np.load(path)  # allow_pickle=False: plain arrays only
    return {k: z[k] for k in z.files}


def as_hits(g, mode):
    n = len(g["rays"])
    h = np.zeros(n, dtype=[("t", "f8"), ("prim", "i8"), ("u", "f8"), ("v", "f8"), ("normal", "f8", 3)])
    h["t"], h["prim"] = g[f"{mode}_t"], g[f"{mode}_prim"]
    h["u"], h["v"] = g[f"{mode}_uv"][:, 0], g[f"{mode}_uv"][:, 1]
    h["normal"] = g[f"{mode}_normal"]
    return h


def test_fixtures_present():
    assert {os.path.basename(f) for f in FILES} == {"rays_cornell.npz", "rays_random_spheres.npz",
                                                      "rays_marched.npz", "rays_marched_f32.npz"}


# The reference's marchers are chaotic in their input: a ray rounded to FP32
# takes other steps, so the FP32 kernel meets them in reference mode on the
# fixture whose rays are exact in FP32 (rays_marched_f32.npz), not on the one
# with FP64 rays.
GPU_CASES = [(path, mode, kernel) for path in FILES for mode in MODES for kernel in ("plain64", "wavefront")
             if not (kernel == "wavefront" and mode == "reference" and os.path.basename(path) == "rays_marched.npz")]


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
@pytest.mark.parametrize("mode", list(MODES))
def test_oracle_reproduces_golden(path, mode):
    import oracle
    g = load(path)
    _, _, world = lm.load_scene(str(g["toml"]))
    h = oracle.trace(world, g["rays"], exact=(mode == "exact"))
    np.testing.assert_array_equal(h["prim"], g[f"{mode}_prim"])
    hit = h["prim"] >= 0
    np.testing.assert_allclose(h["t"][hit], g[f"{mode}_t"][hit], rtol=1e-12)
    np.testing.assert_allclose(h["normal"][hit], g[f"{mode}_normal"][hit], atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("path,mode,kernel", GPU_CASES,
                         ids=[f"{k}-{m}-{os.path.basename(p)}" for p, m, k in GPU_CASES])
def test_gpu_matches_golden(gpu, path, mode, kernel):
    from test_gpu_parity import compare_hits
    g = load(path)
    _, _, world = lm.load_scene(str(g["toml"]))
    ho = as_hits(g, mode)
    hg = lm.trace_rays(world, g["rays"], path=kernel, intersect_mode=MODES[mode])
    r = compare_hits(g["rays"], ho, hg, fp32=(kernel == "wavefront"))
    print(os.path.basename(path), mode, kernel, r)
    assert r["bad_t"] == 0 if kernel == "plain64" else r["bad_t"] <= int(1e-3 * r["n"])
    assert r["flips"] <= (2 if kernel == "plain64" else int(2e-3 * r["n"]) + 1)
