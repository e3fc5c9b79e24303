"""Device tone mapping (grrt_ext_tonemap_device) against the host
grrt_tonemap (src/tonemapper.cpp:7-71) on the same rendered frames.

The device reduction adds the per-pixel ln(1e-6 + L) terms in a tree, the
host in pixel order, and the device pow/log/exp are ROCm's: the exposure can
differ in its last bits, so a mapped value may differ by an ulp of float.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grrt():
    import grrt_amd
    if not grrt_amd.gpu_available():
        pytest.skip("no GPU")
    return grrt_amd


def test_device_tonemap_matches_host(grrt):
    print("case: tone map 320x256 Kerr a=0.998 frame, stars, 4 spp", flush=True)
    p = grrt.make_params(width=320, height=256, spin=0.998, observer_theta=math.radians(80.0),
                         fov=math.radians(90.0), stars=True, spp=4)
    with grrt.Context(p) as ctx:
        host = ctx.render()
        ctx.render_device()
        dev = ctx.get_frame()
        assert np.array_equal(dev, host)
        ms = ctx.tonemap_device()
        got = ctx.get_frame()
    want = grrt.tonemap(host)
    assert ms > 0.0
    assert np.array_equal(got[..., 3], host[..., 3])
    assert np.all((got[..., :3] >= 0) & (got[..., :3] <= 1))
    same = np.mean(got == want)
    diff = np.abs(got.astype(np.float64) - want)
    print(f"identical {same:.6f}, max |diff| {diff.max():.3e}", flush=True)
    assert same >= 0.999
    assert diff.max() <= 4 * 2.0 ** -24


def test_device_tonemap_black_frame_unchanged(grrt):
    print("case: tone map of an all-black frame", flush=True)
    p = grrt.make_params(width=64, height=48, spin=0.9, disk=False, stars=False)
    with grrt.Context(p) as ctx:
        ctx.render_device()
        before = ctx.get_frame()
        ctx.tonemap_device()
        after = ctx.get_frame()
    assert not np.any(before[..., :3])
    assert np.array_equal(after, before)
