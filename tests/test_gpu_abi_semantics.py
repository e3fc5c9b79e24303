"""ABI semantics the reference stubs out (src/api.cpp:185-187, 215-231),
made real here: progress while a frame renders, cancellation from another
thread or from the progress callback, and grrt_update_params rebuilding the
scene.

Cancellation contract: claimed rays always finish, and claims stop once the
flag is seen, so the traced samples are exactly a prefix of the ray queue
(row-major sample order) and every pixel with an untraced sample is zero.
"""
import math
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UNTRACED = 0xFF


def _slow_frame(grrt):
    # 1024^2 x 16 seeded samples: about 70 ms on one MI355X, long enough for
    # the host to see progress and cancel part-way.
    p = grrt.make_params(width=1024, height=1024, spin=0.998, observer_theta=math.radians(80),
                         fov=math.radians(90), stars=True)
    ctx = grrt.Context(p)
    ctx.set_sampling(grrt.SAMPLING_SEEDED, 16, 99)
    return ctx


def _check_prefix(cls, fb, nS):
    traced = cls != UNTRACED
    k = int(traced.sum())
    assert 0 < k < len(cls), k
    assert traced[:k].all() and not traced[k:].any(), "traced samples are not a queue prefix"
    done_px = traced.reshape(-1, nS).all(axis=1)
    px = fb.reshape(-1, 4)
    assert np.all(px[~done_px, :3] == 0.0), "a pixel with an untraced sample is not black"
    assert np.all(px[:, 3] == 1.0)
    return k, done_px


def test_progress_reported_while_rendering(grrt):
    print("case: progress callback and grrt_progress during a 1024^2 x 16 frame", flush=True)
    with _slow_frame(grrt) as ctx:
        ctx.render()                                   # warm
        seen, polled = [], []
        stop = threading.Event()

        def poll():
            while not stop.is_set():
                polled.append(ctx.progress())
                time.sleep(0.0005)

        t = threading.Thread(target=poll)
        t.start()
        fb = ctx.render(progress=seen.append)
        stop.set()
        t.join()
        assert not ctx.cancelled
        assert ctx.progress() == 1.0
    print(f"callback values {len(seen)}: {seen[:3]} .. {seen[-3:]}; polled {len(polled)}", flush=True)
    assert seen[-1] == 1.0
    assert all(b >= a for a, b in zip(seen, seen[1:])), "progress must not fall"
    assert any(0.0 < v < 1.0 for v in seen), "no intermediate progress reported"
    inner = [v for v in polled if 0.0 < v < 1.0]
    assert inner, "grrt_progress never read an intermediate value"
    assert all(b >= a for a, b in zip(inner, inner[1:]))
    assert np.isfinite(fb).all()


def test_cancel_from_progress_callback_leaves_a_queue_prefix(grrt):
    print("case: grrt_cancel from the progress callback, 1024^2 x 16", flush=True)
    with _slow_frame(grrt) as ctx:
        full = ctx.render()
        full_cls = ctx.classes()

        def cb(f):
            if f > 0.2:
                ctx.cancel()

        fb = ctx.render(progress=cb)
        assert ctx.cancelled
        cls = ctx.classes()
        st = ctx.stats()
        assert st.cancelled == 1
        k, done_px = _check_prefix(cls, fb, 16)
        print(f"traced {k} of {len(cls)} samples", flush=True)
        assert np.array_equal(cls[:k], full_cls[:k])
        assert np.array_equal(fb.reshape(-1, 4)[done_px], full.reshape(-1, 4)[done_px])
        # the next call starts afresh: the flag is cleared per call
        again = ctx.render()
        assert not ctx.cancelled and np.array_equal(again, full)


def test_cancel_from_another_thread(grrt):
    print("case: grrt_cancel from a second thread, 1024^2 x 16", flush=True)
    with _slow_frame(grrt) as ctx:
        ctx.render()

        def canceller():
            # grrt_progress still reads the previous call's 1.0 until the
            # render starts: wait for a value inside the render.
            deadline = time.monotonic() + 10.0
            while time.monotonic() < deadline:
                p = ctx.progress()
                if 0.05 < p < 0.9:
                    break
                time.sleep(0.0002)
            ctx.cancel()

        t = threading.Thread(target=canceller)
        t.start()
        fb = ctx.render()
        t.join()
        assert ctx.cancelled
        _check_prefix(ctx.classes(), fb, 16)


def test_cancel_stops_the_fits_stream_between_bands(grrt, tmp_path):
    """A cancel made from the FITS progress callback (called between bands)
    ends the stream with 1; the rows of later bands stay zero in the file."""
    print("case: FITS stream cancelled after its first band, 256^2 x 32 bins", flush=True)
    freqs = 10.0 ** (14 + 4 * np.arange(32) / 31)
    p = grrt.make_params(width=256, height=256, metric=1, spin=0.9, observer_theta=math.radians(75),
                         volumetric=True, disk_outer=20.0, freqs=freqs)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(freqs)
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        cube = ctx.render_spectral()
        calls = []

        def cb(f):
            calls.append(f)
            ctx.cancel()

        rc = ctx.render_spectral_to_fits(str(tmp_path / "c.fits"), progress=cb)
        assert rc == 1
        assert len(calls) == 1
    raw = (tmp_path / "c.fits").read_bytes()
    off = 0
    while not raw[off:off + 80].startswith(b"END "):
        off += 80
    hdr = (off + 80 + 2879) // 2880 * 2880
    data = np.frombuffer(raw, dtype=">f8", count=256 * 256 * 32, offset=hdr).reshape(32, 256, 256)
    img = data[:, ::-1, :].transpose(1, 2, 0)          # FITS stores the bottom row first
    band = max(1, min((1 << 18) // (256 * 32), 256))
    assert np.array_equal(img[:band], cube[:band])
    assert np.all(img[band:] == 0.0)


def test_update_params_equals_fresh_context(grrt):
    print("case: grrt_update_params vs a fresh context", flush=True)
    a = grrt.make_params(width=96, height=64, spin=0.5, observer_theta=1.2, fov=1.0, stars=False)
    b = grrt.make_params(width=80, height=72, spin=0.95, observer_r=30.0, observer_theta=1.45, fov=1.3,
                         stars=True, spp=4, disk_outer=15.0)
    with grrt.Context(b) as fresh:
        want = fresh.render()
    with grrt.Context(a) as ctx:
        ctx.render()
        ctx.update_params(b)
        got = ctx.render()
    assert got.shape == want.shape
    assert np.array_equal(got, want)
