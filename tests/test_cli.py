"""The command line front end (luminara-gpu, src/main.rs) on the CPU: the
error behaviour of the scene argument and the options it accepts.  Its post
chain runs in the library on the GPU (known answers of the chain's
restatement: test_post.py; rendering through the front end:
test_gpu_cli.py)."""
import os
import subprocess

import pytest

from luminara_amd import build

CLI = build.CLI


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        pytest.skip("luminara-gpu is not built (python luminara_amd/build.py)")
    return CLI


def test_missing_and_invalid_scene(cli, tmp_path):
    r = subprocess.run([cli, str(tmp_path / "none.toml")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot read scene file" in r.stderr
    bad = tmp_path / "bad.toml"
    bad.write_text("[render]\nwidth = \n")
    r = subprocess.run([cli, str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "invalid scene" in r.stderr
    # no scene: the reference's demo scene (src/main.rs:198-201); without a GPU
    # the render then fails with a device error
    r = subprocess.run([cli, "-w", "8", "-H", "8", "-s", "1", "-q", "-o", str(tmp_path / "demo.png")],
                       capture_output=True, text=True)
    assert "No scene file provided, rendering demo scene..." in r.stderr
    r = subprocess.run([cli, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--samples" in r.stderr
    for flag in ("--auto-exposure", "--tone-map", "--depth-fog", "--fog-color", "--save-depth", "--save-normals",
                 "--save-albedo", "--save-hdr"):
        assert flag in r.stderr
