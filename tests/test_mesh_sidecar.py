"""`[[mesh]]` PLY ingestion (host C++, no GPU): the text parser's rules and
messages (src/obj.rs:157-246, src/scene.rs:1545-1559) and the binary sidecar
`<file>.lmnmesh` that the loader writes after parsing a PLY file and reads on
later loads while it still describes the file."""
import os
import struct

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi, workloads

MAT = 'material = { type = "lambertian", color = [0.5, 0.5, 0.5] }'


def _toml(path, extra=""):
    return f'[[mesh]]\nfile = "{path}"\n{extra}{MAT}\n'


def _prims(world):
    return np.array([list(p.p)[:9] for p in world.prims])


def _mesh(tmp_path, freq=6, name="m.ply"):
    v, f = workloads.displaced_sphere(freq, [0.25, 1.0, -0.5], 1.3, 11)
    path = os.path.join(str(tmp_path), name)
    workloads.write_ply(path, v, f)
    return path, v, f


def test_sidecar_written_then_read_bit_identical(tmp_path, capfd, monkeypatch):
    monkeypatch.delenv("LMN_MESH_SIDECAR", raising=False)
    path, v, f = _mesh(tmp_path)
    side = path + ".lmnmesh"
    toml = _toml(path, "scale = 1.7\noffset = [0.1, -2.0, 3.0]\n")
    _, _, a = lm.load_scene(toml)
    assert os.path.exists(side)
    first = capfd.readouterr().err
    assert f"Loaded PLY: {len(v)} vertices, {len(f)} faces, {len(f)} triangles\n" in first
    assert "binary sidecar" not in first
    _, _, b = lm.load_scene(toml)
    assert "(binary sidecar)" in capfd.readouterr().err
    pa, pb = _prims(a), _prims(b)
    assert pa.shape == (len(f), 9)
    assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64))  # bit for bit
    # the vertices as the reference computes them: x * scale + offset
    exp = (v * 1.7 + np.array([0.1, -2.0, 3.0]))[f].reshape(-1, 9)
    assert np.array_equal(pa, exp)
    # the sidecar's layout: 64-byte header, FP64 vertices, uint32 indices
    raw = open(side, "rb").read()
    magic, ver, _, nv, nt, nf, size, _, _ = struct.unpack("<8sIIQQQQqQ", raw[:64])
    assert (magic, ver, nv, nt, nf, size) == (b"LMNMESH1", 1, len(v), len(f), len(f), os.path.getsize(path))
    assert len(raw) == 64 + 24 * nv + 12 * nt
    np.testing.assert_array_equal(np.frombuffer(raw, "<f8", 3 * nv, 64).reshape(-1, 3), v)
    np.testing.assert_array_equal(np.frombuffer(raw, "<u4", 3 * nt, 64 + 24 * nv).reshape(-1, 3), f)


def test_stale_sidecar_is_parsed_again(tmp_path, monkeypatch):
    monkeypatch.delenv("LMN_MESH_SIDECAR", raising=False)
    path, v, f = _mesh(tmp_path)
    lm.load_scene(_toml(path))
    st = os.stat(path)
    # same size and modification time, other contents near the start: the
    # digest of the first and last 64 KiB catches it
    text = open(path).read()
    x, y, z = (repr(float(c)) for c in v[0])
    first_vertex = f"{x} {y} {z}"
    moved = f"{x} {y[:-1] + ('1' if y[-1] != '1' else '2')} {z}"
    assert len(moved) == len(first_vertex) and moved != first_vertex
    with open(path, "w") as fh:
        fh.write(text.replace(first_vertex, moved, 1))
    os.utime(path, ns=(st.st_atime_ns, st.st_mtime_ns))
    assert os.path.getsize(path) == st.st_size
    _, _, w = lm.load_scene(_toml(path))
    got = _prims(w)
    v2 = v.copy()
    v2[0, 1] = float(moved.split()[1])
    assert np.array_equal(got, v2[f].reshape(-1, 9))
    # a new modification time alone also invalidates it (and it is rewritten)
    os.utime(path, ns=(st.st_atime_ns, st.st_mtime_ns + 10**9))
    _, _, w = lm.load_scene(_toml(path))
    assert np.array_equal(_prims(w), v2[f].reshape(-1, 9))
    hdr = struct.unpack("<8sIIQQQQqQ", open(path + ".lmnmesh", "rb").read(64))
    assert hdr[7] == st.st_mtime_ns + 10**9


def test_damaged_sidecar_falls_back_to_text(tmp_path, monkeypatch):
    monkeypatch.delenv("LMN_MESH_SIDECAR", raising=False)
    path, v, f = _mesh(tmp_path, freq=4)
    side = path + ".lmnmesh"
    lm.load_scene(_toml(path))
    good = open(side, "rb").read()
    expect = v[f].reshape(-1, 9)
    monkeypatch.setenv("LMN_MESH_SIDECAR", "read")  # read, never rewrite: each case sees its own bytes
    nv = len(v)
    bad_index = bytearray(good)
    bad_index[64 + 24 * nv:64 + 24 * nv + 4] = struct.pack("<I", nv)  # an index past the vertices
    wrong_magic = bytearray(good)
    wrong_magic[0:8] = b"LMNMESH0"
    for damaged in (good[:-4], good + b"\0", bytes(bad_index), bytes(wrong_magic), b""):
        with open(side, "wb") as fh:
            fh.write(damaged)
        _, _, w = lm.load_scene(_toml(path))
        assert np.array_equal(_prims(w), expect)
        assert open(side, "rb").read() == damaged
    # a current sidecar whose vertices were changed is believed: the header
    # vouches for the source, not for the sidecar's own body
    forged = bytearray(good)
    forged[64:72] = struct.pack("<d", 42.0)
    with open(side, "wb") as fh:
        fh.write(forged)
    _, _, w = lm.load_scene(_toml(path))
    assert (_prims(w) == 42.0).any()


def test_sidecar_modes(tmp_path, monkeypatch):
    path, v, f = _mesh(tmp_path, freq=3)
    side = path + ".lmnmesh"
    monkeypatch.setenv("LMN_MESH_SIDECAR", "off")
    lm.load_scene(_toml(path))
    assert not os.path.exists(side)
    monkeypatch.setenv("LMN_MESH_SIDECAR", "read")
    lm.load_scene(_toml(path))
    assert not os.path.exists(side)
    monkeypatch.delenv("LMN_MESH_SIDECAR")
    lm.load_scene(_toml(path))
    assert os.path.exists(side)
    # "off" ignores an existing sidecar, even one that does not match the text
    forged = bytearray(open(side, "rb").read())
    forged[64:72] = struct.pack("<d", 42.0)
    open(side, "wb").write(forged)
    monkeypatch.setenv("LMN_MESH_SIDECAR", "off")
    _, _, w = lm.load_scene(_toml(path))
    assert np.array_equal(_prims(w), v[f].reshape(-1, 9))
    # no sidecar where the directory cannot be written: the load still works
    os.remove(side)
    monkeypatch.delenv("LMN_MESH_SIDECAR")
    os.chmod(str(tmp_path), 0o555)
    try:
        _, _, w = lm.load_scene(_toml(path))
    finally:
        os.chmod(str(tmp_path), 0o755)
    assert np.array_equal(_prims(w), v[f].reshape(-1, 9))


HEADER = "ply\nformat ascii 1.0\nelement vertex {nv}\nproperty float x\nproperty float y\nproperty float z\n" \
         "element face {nf}\nproperty list uchar int vertex_indices\nend_header\n"


@pytest.mark.parametrize("text,msg", [
    ("", "Not a PLY file (missing 'ply' magic)"),
    ("plyx\nend_header\n", "Not a PLY file (missing 'ply' magic)"),
    ("ply\nelement vertex 3\n", "PLY header not terminated (missing 'end_header')"),
    ("ply\nelement vertex 3x\nend_header\n", "Bad vertex count: invalid digit found in string"),
    ("ply\nelement face -1\nend_header\n", "Bad face count: invalid digit found in string"),
    (HEADER.format(nv=2, nf=0) + "0 0 0\n", "Expected vertex 1, got EOF"),
    (HEADER.format(nv=1, nf=0) + "0 0\n", "Vertex 0: need at least 3 coordinates"),
    (HEADER.format(nv=1, nf=0) + "0 0x1 0\n", "Vertex 0 y: invalid float literal"),
    (HEADER.format(nv=1, nf=0) + "0 0 1e\n", "Vertex 0 z: invalid float literal"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n", "Expected face 0, got EOF"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n  \n", "Face 0: empty line"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\nthree 0 1 2\n",
     "Face 0 vertex count: invalid digit found in string"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n3 0 1\n", "Face 0: expected 3 vertex indices"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n4 0 q 1\n", "Face 0: expected 4 vertex indices"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n3 0 q 1\n", "Face 0 index 2: invalid digit found in string"),
    (HEADER.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n", "Face 0: vertex index out of bounds"),
])
def test_ply_messages(tmp_path, text, msg, monkeypatch):
    monkeypatch.setenv("LMN_MESH_SIDECAR", "off")
    path = tmp_path / "e.ply"
    path.write_text(text)
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene(_toml(path))
    assert e.value.code == _abi.LMN_ERR_PARSE
    assert str(e.value).endswith(msg), str(e.value)


def test_ply_accepts_what_the_reference_accepts(tmp_path, monkeypatch):
    monkeypatch.setenv("LMN_MESH_SIDECAR", "off")
    # CRLF lines, tabs, '+' signs, extra tokens, a pentagon fan, degenerate
    # faces with fewer than three indices, trailing lines after the faces
    text = ("ply\r\nformat ascii 1.0\r\nelement vertex +5\r\ncomment x\r\nelement face 3\r\n end_header \r\n"
            "0 0 0 9 9\r\n1\t0 0\r\n1 1 0\r\n0.5 1.5 0\r\n-0 1 +0e0\r\n"
            "5 0 1 2 3 4 77\r\n2 0 1\r\n+3 +4 3 2\r\ntrailing words\r\n")
    path = tmp_path / "ok.ply"
    path.write_text(text)
    _, _, w = lm.load_scene(_toml(path))
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0.5, 1.5, 0], [0, 1, 0]], float)
    tris = [(0, 1, 2), (0, 2, 3), (0, 3, 4), (4, 3, 2)]
    assert np.array_equal(_prims(w), v[np.array(tris)].reshape(-1, 9))


def test_missing_mesh_message(tmp_path):
    with pytest.raises(_abi.LuminaraError) as e:
        lm.load_scene(_toml(tmp_path / "missing.ply"))
    assert e.value.code == _abi.LMN_ERR_IO
    assert str(e.value).endswith("missing.ply': No such file or directory (os error 2)")
