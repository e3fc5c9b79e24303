"""Row-streamed HDR output (grrt_ext_render_hdr): the CLI's linear and
exposure-scaled Radiance files (reference cli/main.cpp:398-434) written
while the frame renders in bands of rows.

The checks: the frame equals grrt_render's bit for bit; the streamed linear
file and the scaled file equal, byte for byte, the whole-frame writer's
output for that frame (grrt_ext_write_hdr); progress is reported once per
band and ends at 1; a cancel from the progress callback after the first band
returns 1 with the unrendered rows zero, a complete linear file (black rows
after the cancel point) and no scaled file."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 1024, 700            # bands of 2^18 / 1024 = 256 rows: 256, 256, 188


def _params(grrt):
    return grrt.make_params(width=W, height=H, metric=grrt.METRIC_KERR, spin=0.998,
                            observer_theta=math.radians(80.0), fov=math.radians(90.0), stars=True)


def test_streamed_hdr_equals_whole_frame_files(grrt, tmp_path):
    print("case: streamed HDR 1024x700 Kerr a=0.998 with stars", flush=True)
    lin, scaled = str(tmp_path / "f_linear.hdr"), str(tmp_path / "f.hdr")
    seen = []
    with grrt.Context(_params(grrt)) as ctx:
        rc, fb = ctx.render_hdr(lin, scaled, progress=seen.append)
        ref = ctx.render()
    assert rc == 0
    assert np.array_equal(fb, ref)
    assert np.any(fb[..., :3] > 0)
    assert seen == pytest.approx([256 / H, 512 / H, 1.0])
    lin_ref, scaled_ref = str(tmp_path / "r_linear.hdr"), str(tmp_path / "r.hdr")
    grrt.write_hdr(lin_ref, ref)
    grrt.write_hdr(scaled_ref, ref, normalised=True)
    with open(lin, "rb") as a, open(lin_ref, "rb") as b:
        assert a.read() == b.read()
    with open(scaled, "rb") as a, open(scaled_ref, "rb") as b:
        assert a.read() == b.read()


def test_streamed_hdr_cancel_after_first_band(grrt, tmp_path):
    print("case: streamed HDR cancelled from the progress callback", flush=True)
    lin, scaled = str(tmp_path / "c_linear.hdr"), str(tmp_path / "c.hdr")
    with grrt.Context(_params(grrt)) as ctx:
        full = ctx.render()
        rc, fb = ctx.render_hdr(lin, scaled, progress=lambda f: ctx.cancel())
    assert rc == 1
    assert np.array_equal(fb[:256], full[:256])
    assert not np.any(fb[256:])
    assert not os.path.exists(scaled)
    lin_ref = str(tmp_path / "c_ref.hdr")
    grrt.write_hdr(lin_ref, fb)
    with open(lin, "rb") as a, open(lin_ref, "rb") as b:
        assert a.read() == b.read()
