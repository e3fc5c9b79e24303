"""grrt_ext_render_device_noalloc: the frame of grrt_ext_render_device with
no device allocation, its launch block, counters and per-sample buffers in
the library's fixed per-GPU scratch array.  Checked through the C ABI
against grrt_ext_render_device on two calls in a row, the second with a
smaller frame (the scratch keeps the first call's contents), and for the
size limit."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipDeviceSynchronize.argtypes = []
    return L


def _frame_pair(grrt, hip, p, sampling):
    """(no-alloc frame, classes), (grrt_ext_render_device frame, classes)."""
    n = p.width * p.height * 4
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), n * 4) == 0
    try:
        with grrt.Context(p) as ctx:
            ctx.set_sampling(*sampling)
            assert hip.hipMemset(buf, 0x7F, n * 4) == 0       # stale bytes the call must overwrite
            rc = grrt.lib().grrt_ext_render_device_noalloc(ctx._h, buf)
            assert rc == 0, ctx.error
            a = np.zeros((p.height, p.width, 4), np.float32)
            assert hip.hipMemcpy(a.ctypes.data, buf, n * 4, 2) == 0
            ca = ctx.classes()
            st = ctx.stats()
            assert st.rays == ctx.num_samples()
            ctx.render_device()                                # the existing entry point, the context's frame
            b = ctx.get_frame()
            cb = ctx.classes()
    finally:
        hip.hipFree(buf)
    return (a, ca), (b, cb)


def test_noalloc_equals_render_device_twice(grrt, hip):
    print("case: no-alloc vs render_device, Kerr a=0.998 1024^2 x 4 seeded, then 96x64 centres a=0.5", flush=True)
    big = grrt.make_params(width=1024, height=1024, spin=0.998, observer_theta=math.radians(80),
                           fov=math.radians(90), stars=True)
    (a, ca), (b, cb) = _frame_pair(grrt, hip, big, (grrt.SAMPLING_SEEDED, 4, 1234))
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    assert np.isfinite(a).all() and a[..., :3].max() > 0
    small = grrt.make_params(width=96, height=64, spin=0.5, observer_theta=1.2, fov=1.0, stars=False)
    (a2, ca2), (b2, cb2) = _frame_pair(grrt, hip, small, (grrt.SAMPLING_CENTER, 1, 0))
    assert np.array_equal(a2, b2) and np.array_equal(ca2, cb2)
    assert len(ca2) == 96 * 64


def test_noalloc_rejects_frames_beyond_its_scratch(grrt, hip):
    p = grrt.make_params(width=2048, height=2048, spin=0.998)
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), 2048 * 2048 * 16) == 0
    try:
        with grrt.Context(p) as ctx:
            ctx.set_sampling(grrt.SAMPLING_SEEDED, 4, 1)
            rc = grrt.lib().grrt_ext_render_device_noalloc(ctx._h, buf)
            assert rc == -1
            assert "MiB of scratch" in ctx.error
            assert grrt.lib().grrt_ext_render_device_noalloc(ctx._h, None) == -1
    finally:
        hip.hipFree(buf)
