"""luminara-gpu renders through the library and writes the reference's
outputs: the PNG/PPM of the default post chain (exposure, ACES, sRGB,
8-bit) of the same framebuffer `render` returns, the Radiance HDR of the
exposed framebuffer, and the statistics JSON (src/main.rs:774-989)."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import build, workloads
import oracle

pytestmark = pytest.mark.gpu


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 6
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, w * 4 + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4)


def read_ppm(path):
    tok = open(path).read().split()
    assert tok[0] == "P3"
    w, h = int(tok[1]), int(tok[2])
    return np.array(tok[4:], dtype=np.int64).reshape(h, w, 3)


def read_rgbe(path):
    data = open(path, "rb").read()
    head, body = data.split(b"\n\n", 1)
    line, pix = body.split(b"\n", 1)
    _, h, _, w = line.split()
    e = np.frombuffer(pix, dtype=np.uint8).reshape(int(h), int(w), 4).astype(np.float64)
    scale = np.where(e[..., 3:] > 0, np.ldexp(1.0, (e[..., 3:] - 128).astype(int)) / 256.0, 0.0)
    return e[..., :3] * scale


def test_cli_outputs_match_the_library_render(gpu, tmp_path):
    scene = tmp_path / "cornell.toml"
    scene.write_text(workloads.cornell_box(64, 64, 16, 8))
    args = [build.CLI, str(scene), "-w", "48", "-H", "40", "-s", "9", "-d", "6", "--seed", "7", "-q"]
    png, ppm, hdr, js = (tmp_path / n for n in ("o.png", "o.ppm", "o.hdr", "o.json"))
    subprocess.run(args + ["-o", str(png), "--save-hdr", str(hdr), "--save-json", str(js)], check=True, timeout=120)
    subprocess.run(args + ["-o", str(ppm)], check=True, timeout=120)
    cfg, cam, world = lm.load_scene(scene.read_text())
    cfg.width, cfg.height, cfg.samples_per_pixel, cfg.max_depth, cfg.seed = 48, 40, 9, 6, 7
    rows = lm.render(cfg, cam, world).rows
    img = read_png(png)
    assert img.shape == (40, 48, 4) and np.all(img[..., 3] == 255)
    expect, exposed, _ = oracle.post(rows, exposure=cfg.exposure, tone_map=cfg.tone_map)
    # the GPU's and the host's pow may round the last bit apart
    assert np.abs(img[..., :3].astype(np.int64) - expect[..., :3]).max() <= 1
    assert (img[..., :3] == expect[..., :3]).mean() > 0.999
    assert np.array_equal(read_ppm(ppm), img[..., :3].astype(np.int64))
    raw = open(hdr, "rb").read()
    assert raw.startswith(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 40 +X 48\n")
    assert np.array_equal(np.frombuffer(raw.split(b"+X 48\n", 1)[1], np.uint8).reshape(40, 48, 4),
                          oracle.rgbe(exposed))
    h = read_rgbe(hdr)
    ref = rows.astype(np.float64) * cfg.exposure
    m = ref.max(axis=2, keepdims=True)
    assert np.all(np.abs(h - ref) <= m / 128.0 + 1e-30)
    st = json.loads(js.read_text())
    assert st["width"] == 48 and st["height"] == 40 and st["samples"] == 9 and st["max_depth"] == 6
    assert st["total_rays"] == 48 * 40 * 9


def test_cli_preview_and_benchmark(gpu, tmp_path):
    scene = tmp_path / "cornell.toml"
    scene.write_text(workloads.cornell_box(64, 64, 16, 8))
    png = tmp_path / "p.png"
    subprocess.run([build.CLI, str(scene), "-p", "-q", "-o", str(png)], check=True, timeout=120)
    assert read_png(png).shape == (16, 16, 4)
    # the benchmark renders the demo scene at 400x225, 64 spp, whatever scene is given (src/main.rs:688-706)
    r = subprocess.run([build.CLI, str(scene), "--benchmark", "-s", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Benchmark result:" in r.stderr and "Mrays/s" in r.stderr
    assert "rendering demo scene at 400x225 @ 64 spp" in r.stderr and "5760000 primary rays" in r.stderr


def test_cli_demo_scene(gpu, tmp_path):
    png = tmp_path / "demo.png"
    r = subprocess.run([build.CLI, "-w", "80", "-H", "45", "-s", "4", "-q", "-o", str(png)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "No scene file provided, rendering demo scene..." in r.stderr
    img = read_png(png)
    assert img.shape == (45, 80, 4) and img[..., :3].std() > 10   # sky, ground and spheres


def test_cli_passes_fog_and_tone_flags(gpu, tmp_path):
    scene = tmp_path / "cornell.toml"
    scene.write_text(workloads.cornell_box(64, 64, 16, 8))
    dp, npth, ap, png = (tmp_path / n for n in ("d.png", "n.png", "a.png", "o.png"))
    args = [build.CLI, str(scene), "-w", "40", "-H", "40", "-s", "4", "-d", "4", "--seed", "3", "-q",
            "--tone-map", "reinhard", "--depth-fog", "0.001", "--fog-color", "0.2,0.3,0.4", "-o", str(png),
            "--save-depth", str(dp), "--save-normals", str(npth), "--save-albedo", str(ap)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Saved depth pass to" in r.stderr and "Saved albedo pass to" in r.stderr
    cfg, cam, world = lm.load_scene(scene.read_text())
    cfg.width, cfg.height, cfg.samples_per_pixel, cfg.max_depth, cfg.seed = 40, 40, 4, 4, 3
    cfg.tone_map, cfg.depth_fog, cfg.depth_fog_color = 1, 0.001, (0.2, 0.3, 0.4)
    img = lm.render_image(cfg, world, save_depth=True, save_normals=True, save_albedo=True)
    assert np.array_equal(read_png(png), img.pixels)
    assert np.array_equal(read_png(npth)[..., :3], img.normal_pass)
    assert np.array_equal(read_png(ap)[..., :3], img.albedo_pass)
    # depth image (src/main.rs:856-876): depth * 255 / max depth, truncated
    d = img.depth_pass
    norm = np.float32(255.0) / d.max() if d.max() > 0 else np.float32(1.0)
    want = np.clip(d * norm, 0, 255).astype(np.uint8)
    assert np.array_equal(read_png(dp)[..., 0], want)
