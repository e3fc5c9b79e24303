"""The texture background (GRRT_BG_TEXTURE with background_texture_path).

The reference declares the mode (include/grrt/types.h:15-17) and renders it
black: its scene builds a celestial sphere for GRRT_BG_STARS only
(src/api.cpp:136).  Here a path selects an equirectangular image, sampled
bilinearly at the escape direction; with no path the frame stays the
reference's black one.

The sampling convention is pinned by a smooth map of the direction
(parity_util.direction_rgb, evaluated at texel centres): bilinear
interpolation of a function with second derivatives <= 1/2 per rad^2 on
a 1024 x 512 map errs by at most (h_theta^2 + h_phi^2) / 16 ~ 5e-6
(h = pi/512), so the tolerance 1e-5 holds it, while a half-texel offset, a
flipped row order or a phi shift errs by > 1e-3.  Rays within two texel
heights of a pole are left out: there the map is clamped to its first or
last row.  Non-escaped rays must be untouched, bit for bit.
"""
import math

import numpy as np
import pytest

from parity_util import direction_rgb, direction_texture, write_pfm

pytestmark = pytest.mark.gpu

TEX_W, TEX_H = 1024, 512
TOL = 1e-5


@pytest.fixture(scope="module")
def dir_map(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("tex") / "direction.pfm")
    write_pfm(path, direction_texture(TEX_W, TEX_H))
    return path


def _render(grrt, p):
    with grrt.Context(p) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.set_records(True)
        fb = ctx.render()
        rec = ctx.records()
    return fb, rec


def _check_escaped(rec_tex, rec_black):
    esc = rec_tex["term"] == 1
    assert np.array_equal(esc, rec_black["term"] == 1)
    assert np.array_equal(rec_tex["color"][~esc], rec_black["color"][~esc])
    th, ph = rec_tex["x_final"][esc, 2], rec_tex["x_final"][esc, 3]
    away = np.abs(np.sin(th)) > math.sin(2.0 * math.pi / TEX_H)
    got = rec_tex["color"][esc] - rec_black["color"][esc]
    err = np.abs(got - direction_rgb(th, ph))[away]
    print(f"escaped rays {esc.sum()}, away from the poles {away.sum()}, max error {err.max():.2e}", flush=True)
    assert away.sum() > 100
    assert err.max() <= TOL


def test_thin_disk_texture_background(grrt, dir_map):
    print("case: thin disk Kerr a=0.9 128^2, direction map 1024x512", flush=True)
    kw = dict(width=128, height=128, metric=1, spin=0.9, observer_theta=math.radians(75.0),
              fov=math.radians(90.0), disk=True, disk_outer=20.0)
    fb_t, rec_t = _render(grrt, grrt.make_params(texture=dir_map, **kw))
    fb_b, rec_b = _render(grrt, grrt.make_params(**kw))
    _check_escaped(rec_t, rec_b)
    # the frame is the records' colour (one centred sample per pixel)
    assert np.array_equal(fb_t[..., :3].reshape(-1, 3), rec_t["color"].astype(np.float32))


def test_volumetric_rgb_texture_background(grrt, dir_map):
    print("case: volumetric RGB a=0.9 48^2, direction map 1024x512", flush=True)
    kw = dict(width=48, height=48, metric=1, spin=0.9, observer_theta=math.radians(75.0),
              fov=math.radians(90.0), volumetric=True, disk_outer=20.0)
    _, rec_t = _render(grrt, grrt.make_params(texture=dir_map, **kw))
    _, rec_b = _render(grrt, grrt.make_params(**kw))
    _check_escaped(rec_t, rec_b)


def test_no_path_is_the_reference_black_background(grrt):
    print("case: GRRT_BG_TEXTURE without a path, 64^2", flush=True)
    kw = dict(width=64, height=64, metric=1, spin=0.5, disk=True)
    p = grrt.make_params(**kw)
    p.background_type = grrt.BG_TEXTURE
    fb_t, _ = _render(grrt, p)
    fb_b, _ = _render(grrt, grrt.make_params(**kw))
    assert np.array_equal(fb_t, fb_b)


def test_update_params_switches_the_background(grrt, dir_map, tmp_path):
    print("case: update_params texture -> black -> constant texture, 64^2", flush=True)
    kw = dict(width=64, height=64, metric=0, spin=0.0, disk=True)
    const = str(tmp_path / "const.pfm")
    write_pfm(const, np.broadcast_to(np.float32([0.25, 0.5, 1.0]), (4, 8, 3)))
    with grrt.Context(grrt.make_params(texture=dir_map, **kw)) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.set_records(True)
        ctx.render()
        r0 = ctx.records()
        esc = (r0["term"] == 1) & (r0["cls"] == grrt.CLASS_ESCAPED)     # escaped without crossing the disk
        assert esc.any()
        ctx.update_params(grrt.make_params(**kw))
        black = ctx.render()
        rec = ctx.records()
        assert np.all(rec["color"][esc] == 0.0)
        ctx.update_params(grrt.make_params(texture=const, **kw))
        ctx.render()
        rec2 = ctx.records()
    # a constant map returns its colour exactly (in float) for every escaped ray
    assert np.array_equal(rec2["color"][esc].astype(np.float32), np.broadcast_to(np.float32([0.25, 0.5, 1.0]),
                                                                                 (esc.sum(), 3)))
    out = r0["term"] == 1
    assert np.array_equal(rec2["color"][~out], rec["color"][~out])
    crossed = out & ~esc          # crossed the disk, then escaped: disk light plus the map
    assert np.allclose(rec2["color"][crossed] - rec["color"][crossed], [0.25, 0.5, 1.0], rtol=0, atol=1e-9)
    assert np.isfinite(black).all()


def test_unreadable_texture_fails_create(grrt, tmp_path):
    print("case: grrt_create with a missing and a malformed texture", flush=True)
    missing = str(tmp_path / "nope.hdr")
    with pytest.raises(RuntimeError, match="nope.hdr: cannot open"):
        grrt.Context(grrt.make_params(width=16, height=16, texture=missing))
    bad = tmp_path / "bad.hdr"
    bad.write_bytes(b"#?RADIANCE\n\n+Y 2 +X 2\n")
    with pytest.raises(RuntimeError, match="resolution line"):
        grrt.Context(grrt.make_params(width=16, height=16, texture=str(bad)))


def _const_map(tmp_path, colour):
    path = str(tmp_path / "const.pfm")
    write_pfm(path, np.broadcast_to(np.float32(colour), (4, 8, 3)))
    return path


def test_batch_views_take_the_texture(grrt, tmp_path):
    """grrt_ext_render_batch: every escaped pixel of every view carries the
    constant map's colour, captured pixels equal the batch without it, and
    disk pixels gain the colour exactly where their ray escaped afterwards."""
    print("case: batch of 3 views with a constant texture", flush=True)
    colour = [0.25, 0.5, 1.0]
    views = [(0.998, 50.0, math.radians(80.0), 0.0, math.radians(90.0), 96, 64),
             (0.0, 30.0, math.radians(60.0), 0.0, math.radians(70.0), 40, 72),
             (-0.5, 80.0, math.radians(88.0), 0.3, math.radians(40.0), 128, 48)]
    kw = dict(width=64, height=64, metric=1, spin=0.5, disk=True, disk_outer=20.0)
    out = {}
    for tex in (_const_map(tmp_path, colour), None):
        with grrt.Context(grrt.make_params(texture=tex, **kw)) as ctx:
            ctx.set_sampling(grrt.SAMPLING_CENTER)
            frames = ctx.render_batch(views)
            out[tex is None] = (np.concatenate([f[..., :3].reshape(-1, 3) for f in frames]), ctx.classes())
    (rgb_t, cls_t), (rgb_b, cls_b) = out[False], out[True]
    assert np.array_equal(cls_t, cls_b)
    esc = cls_t == grrt.CLASS_ESCAPED
    print(f"escaped pixels {esc.sum()} of {esc.size}", flush=True)
    assert esc.sum() > 1000
    assert np.array_equal(rgb_t[esc], np.broadcast_to(np.float32(colour), (esc.sum(), 3)))
    cap = cls_t == grrt.CLASS_CAPTURED
    assert np.array_equal(rgb_t[cap], rgb_b[cap])
    # disk pixels: the disk's light, plus the map's colour where the ray
    # escaped after crossing the disk (to the float rounding of the sum)
    disk = cls_t == grrt.CLASS_DISK
    d = rgb_t[disk].astype(np.float64) - rgb_b[disk]
    tol = 4e-7 * np.abs(rgb_t[disk]).astype(np.float64) + 1e-6
    plain = np.all(np.abs(d) <= tol, 1)
    lit = np.all(np.abs(d - colour) <= tol, 1)
    print(f"disk pixels {disk.sum()}: unchanged {plain.sum()}, plus the map {lit.sum()}", flush=True)
    assert np.all(plain | lit) and lit.any()


def test_samples_per_pixel_average_the_texture(grrt, tmp_path):
    """Four seeded samples per pixel: every escaped sample carries the map's
    colour, and a pixel whose four samples all escaped is exactly that colour."""
    print("case: 64^2 x 4 spp seeded, constant texture", flush=True)
    colour = [0.125, 0.75, 2.0]
    p = grrt.make_params(width=64, height=64, metric=1, spin=0.9, disk=True, texture=_const_map(tmp_path, colour))
    with grrt.Context(p) as ctx:
        ctx.set_sampling(grrt.SAMPLING_SEEDED, 4, 11)
        fb = ctx.render()
        cls = ctx.classes()
        rgb = ctx.samples()
    esc = cls == grrt.CLASS_ESCAPED
    assert np.array_equal(rgb[esc].astype(np.float32), np.broadcast_to(np.float32(colour), (esc.sum(), 3)))
    all_esc = esc.reshape(64 * 64, 4).all(1)
    assert all_esc.sum() > 500
    assert np.array_equal(fb[..., :3].reshape(-1, 3)[all_esc], np.broadcast_to(np.float32(colour), (all_esc.sum(), 3)))
