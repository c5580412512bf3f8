"""lmn_post_device_noalloc (scratch in a static device array) against
lmn_post_device (scratch allocated by the library), through the C ABI: the
same bytes, exposed HDR and exposure, on consecutive calls whose second frame
is smaller than the first (the static scratch keeps the first call's
contents), with and without the automatic exposure."""
import ctypes as C

import numpy as np
import pytest

import luminara_amd as lm
from luminara_amd import _abi
from luminara_amd._abi import check, lib

from hipmem import DeviceBuffer

pytestmark = pytest.mark.gpu


def _frame(rng, h, w):
    hdr = (rng.random((h, w, 3)) ** 3 * 8.0).astype(np.float32)
    hdr[rng.random((h, w)) < 0.01] = 0.0
    depth = np.where(rng.random((h, w)) < 0.2, np.inf, rng.random((h, w)) * 30.0).astype(np.float32)
    return hdr, depth


def _run(fn, hdr, depth, post):
    h, w, _ = hdr.shape
    n = h * w
    d_hdr, d_dep = DeviceBuffer.from_array(hdr.reshape(-1)), DeviceBuffer.from_array(depth.reshape(-1))
    d_rgba, d_hx = DeviceBuffer(n * 4), DeviceBuffer(n * 12)
    ex = C.c_double(-1.0)
    print(f"[no-alloc] {fn} {w}x{h} auto={post.auto_exposure} fog={post.depth_fog}", flush=True)
    check(getattr(lib(), fn)(C.c_void_p(d_hdr.addr()), w, h, C.byref(post), C.c_void_p(d_dep.addr()),
                             C.c_void_p(d_rgba.addr()), C.c_void_p(d_hx.addr()), C.byref(ex), None))
    return d_rgba.to_array(np.uint8, (n, 4)), d_hx.to_array(np.float32, (n, 3)), ex.value


@pytest.mark.parametrize("auto", [True, False])
def test_noalloc_post_equals_post_device(gpu, auto):
    rng = np.random.default_rng(11)
    cfg = lm.RenderConfig(exposure=0.8, tone_map=0, auto_exposure=auto, depth_fog=0.05,
                          depth_fog_color=(0.7, 0.8, 0.95))
    post = lm.post_params(cfg)
    for h, w in [(270, 480), (97, 131)]:  # the second frame is smaller
        hdr, depth = _frame(rng, h, w)
        a_rgba, a_hx, a_ex = _run("lmn_post_device", hdr, depth, post)
        b_rgba, b_hx, b_ex = _run("lmn_post_device_noalloc", hdr, depth, post)
        assert a_ex == b_ex
        if auto:
            assert a_ex != 0.8  # the automatic exposure was applied
        assert np.array_equal(a_rgba, b_rgba)
        assert np.array_equal(a_hx, b_hx)
        assert (a_rgba[:, 3] == 255).all()


def test_noalloc_post_rejects_bad_arguments(gpu):
    post = lm.post_params(lm.RenderConfig(auto_exposure=True))
    rc = lib().lmn_post_device_noalloc(None, 4, 4, C.byref(post), None, None, None, None, None)
    assert rc == _abi.LMN_ERR_INVALID
    assert b"bad argument" in lib().lmn_last_error()
