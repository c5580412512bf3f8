"""The command line front end (luminara-gpu, src/main.rs) on the CPU: the
default post chain against the reference's own known answers for its tone
maps and sRGB quantisation (src/render.rs:3665-3772) and closed forms, and
the error behaviour of the scene argument.  Rendering through it is a GPU
test (test_gpu_cli.py)."""
import os
import subprocess

import numpy as np
import pytest

from luminara_amd import build

CLI = build.CLI


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        pytest.skip("luminara-gpu is not built (python luminara_amd/build.py)")
    return CLI


def post(cli, values):
    out = subprocess.run([cli, "--post-check", *[repr(float(v)) for v in values]], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    rows = [list(map(float, l.split())) for l in out if l.strip()]
    a = np.array(rows)
    return a[:, 1], a[:, 2], a[:, 3], a[:, 4].astype(int)


def test_tone_maps_reference_known_answers(cli):
    aces, rein, film, _ = post(cli, [0.0, 1.0])
    assert aces[0] == 0.0                     # aces_tonemap_zero
    assert rein[0] == 0.0 and abs(rein[1] - 0.5) < 1e-6   # reinhard_tonemap_basic
    assert abs(film[0]) < 0.01                # filmic_tonemap_zero
    # closed form of the ACES fit at 1: (2.51 + 0.03) / (2.43 + 0.59 + 0.14)
    assert aces[1] == pytest.approx(2.54 / 3.16, rel=1e-15)
    xs = np.arange(100) * 0.5                 # *_tonemap_bounded
    aces, rein, film, _ = post(cli, xs)
    for y in (aces, film):
        assert np.all((y >= 0.0) & (y <= 1.0))
    xs = np.arange(1, 100) * 0.1              # *_tonemap_monotonic
    aces, rein, film, _ = post(cli, xs)
    for y in (aces, rein, film):
        assert np.all(np.diff(y) >= 0.0)
    _, rein, _, _ = post(cli, np.arange(1, 50))
    assert np.all(np.diff(rein) > 0) and np.all(rein < 1.0)


def test_srgb_quantisation_reference_known_answers(cli):
    _, _, _, b = post(cli, [0.0, 1.0, -1.0, 0.5, 0.0031308, 2.0])
    assert b[0] == 0 and b[1] == 255 and b[2] == 0      # srgb_black / _white / _clamps_negative
    # 1.055 * 0.5^(1/2.4) - 0.055 = 0.73536; floor(0.73536 * 256) = 188
    assert b[3] == 188
    # linear segment: 12.92 * 0.0031308 = 0.04045; floor(0.04045 * 256) = 10
    assert b[4] == 10 and b[5] == 255
    _, _, _, b = post(cli, np.arange(1, 101) / 100.0)   # srgb_monotonic
    assert np.all(np.diff(b) >= 0)


def test_missing_and_invalid_scene(cli, tmp_path):
    r = subprocess.run([cli, str(tmp_path / "none.toml")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot read scene file" in r.stderr
    bad = tmp_path / "bad.toml"
    bad.write_text("[render]\nwidth = \n")
    r = subprocess.run([cli, str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "invalid scene" in r.stderr
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "no scene file" in r.stderr
    r = subprocess.run([cli, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--samples" in r.stderr
