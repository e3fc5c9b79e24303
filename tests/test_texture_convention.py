"""The texture background's sampling convention on the CPU: that the GPU
test's tolerance (1e-5 against the smooth direction map) holds the correct
convention and rejects the plausible mistakes.

The sampler below restates the convention the tracer documents (texel
centres at theta = (j + 1/2) pi / H, phi = -pi + (i + 1/2) 2 pi / W,
bilinear, wrapping in phi, clamped at the poles); each planted mistake
changes one detail of it.  No GPU and no library call.
"""
import math

import numpy as np
import pytest

from parity_util import direction_rgb, direction_texture

W, H = 1024, 512
TOL = 1e-5


def _sample(tex, theta, phi, dx=0.0, dy=0.0, flip=False, phi_shift=0.0):
    h, w, _ = tex.shape
    theta = np.mod(theta, 2 * np.pi)
    back = theta > np.pi
    theta = np.where(back, 2 * np.pi - theta, theta)
    phi = np.where(back, phi + np.pi, phi) + phi_shift
    phi = np.mod(phi + np.pi, 2 * np.pi) - np.pi
    x = (phi + np.pi) / (2 * np.pi) * w - 0.5 + dx
    y = theta / np.pi * h - 0.5 + dy
    if flip:
        y = h - 1 - y
    x0 = np.floor(x)
    tx = x - x0
    x0 = x0.astype(int) % w
    x1 = (x0 + 1) % w
    y0 = np.floor(y)
    ty = y - y0
    y0 = y0.astype(int)
    lo, hi = y0 < 0, y0 + 1 > h - 1
    ty = np.where(lo | hi, 0.0, ty)
    y0 = np.where(lo, 0, np.where(hi, h - 1, y0))
    y1 = np.where(lo | hi, y0, y0 + 1)
    a = (1 - tx)[:, None] * tex[y0, x0] + tx[:, None] * tex[y0, x1]
    b = (1 - tx)[:, None] * tex[y1, x0] + tx[:, None] * tex[y1, x1]
    return (1 - ty)[:, None] * a + ty[:, None] * b


@pytest.fixture(scope="module")
def dirs():
    rng = np.random.default_rng(11)
    n = 200_000
    th = np.arccos(rng.uniform(-1, 1, n))            # uniform on the sphere
    ph = rng.uniform(-np.pi, np.pi, n)
    # raw coordinates past the pole, as integrated geodesics leave them
    wrap = rng.uniform(size=n) < 0.3
    th = np.where(wrap, 2 * np.pi - th, th)
    ph = np.where(wrap, ph - np.pi, ph) + 2 * np.pi * rng.integers(-2, 3, n)
    away = np.abs(np.sin(th)) > math.sin(2 * math.pi / H)
    return th[away], ph[away]


@pytest.fixture(scope="module")
def tex():
    return direction_texture(W, H).astype(np.float64)


def test_correct_convention_is_within_the_interpolation_bound(tex, dirs):
    th, ph = dirs
    err = np.abs(_sample(tex, th, ph) - direction_rgb(th, ph)).max()
    h = math.pi / H
    assert err <= 2 * h * h / 16 + 1e-7 <= TOL, err


@pytest.mark.parametrize("mistake", [dict(dx=0.5), dict(dy=0.5), dict(dx=-0.5), dict(flip=True),
                                     dict(phi_shift=math.pi), dict(phi_shift=2 * math.pi / W)])
def test_planted_convention_mistakes_exceed_the_tolerance(tex, dirs, mistake):
    th, ph = dirs
    err = np.abs(_sample(tex, th, ph, **mistake) - direction_rgb(th, ph)).max()
    assert err > 100 * TOL, (mistake, err)
