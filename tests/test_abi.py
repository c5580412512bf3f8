"""The C-ABI library loads and exports every symbol include/luminara.h declares."""
import ctypes
import os
import re

from luminara_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "luminara.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lmn_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_header_symbols_exported():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in luminara.h but not exported"


def test_every_declared_symbol_has_a_binding():
    assert set(declared_functions()) == set(_abi.SIGNATURES)


def test_struct_sizes_match_c_layout():
    import ctypes as C
    assert C.sizeof(_abi.Prim) == 112
    assert C.sizeof(_abi.Transform) == 200
    assert C.sizeof(_abi.Material) == 80
    assert C.sizeof(_abi.Light) == 136
    assert C.sizeof(_abi.Background) == 88
    assert C.sizeof(_abi.Camera) == 152
    assert C.sizeof(_abi.RenderParams) == 104
    assert C.sizeof(_abi.Ray) == 64
    assert C.sizeof(_abi.Hit) == 64
    assert C.sizeof(_abi.RenderConfig) == 168
    assert C.sizeof(_abi.PostParams) == 48
    assert C.sizeof(_abi.Image) == 72
    assert C.sizeof(_abi.Stats) == 104
    assert C.sizeof(_abi.TraceCounts) == 48
    assert C.sizeof(_abi.VertexAttr) == 128
