"""The C-ABI library loads, exports every symbol its headers declare, keeps
the reference's GRRTParams layout, and fails loudly without a GPU.  No
compute calls are made here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "grrt", "api.h"), os.path.join(ROOT, "include", "grrt_ext.h")]


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        names += re.findall(r"GRRT_EXPORT[^;(]*?\b(grrt_\w+)\s*\(", text)
    return names


def test_headers_declare_reference_abi():
    names = set(declared_symbols())
    reference_16 = {"grrt_create", "grrt_destroy", "grrt_update_params", "grrt_render",
                    "grrt_render_cb", "grrt_render_tile", "grrt_cancel", "grrt_progress",
                    "grrt_error", "grrt_cuda_available", "grrt_last_error", "grrt_tonemap",
                    "grrt_debug_pixel", "grrt_set_frequency_bins", "grrt_render_spectral",
                    "grrt_render_spectral_cb", "grrt_render_spectral_to_fits_cb"}
    assert reference_16 <= names


def test_library_exports_every_declared_symbol():
    import grrt_amd
    L = grrt_amd.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    bound = {n for n, _, _ in grrt_amd.API}
    assert set(declared_symbols()) == bound


def test_params_layout_matches_reference():
    import grrt_amd
    P = grrt_amd.GRRTParams
    assert C.sizeof(P) == 240
    offs = {name: getattr(P, name).offset for name, _ in P._fields_}
    expect = {"width": 0, "height": 4, "metric_type": 8, "mass": 16, "spin": 24, "observer_r": 32,
              "observer_theta": 40, "observer_phi": 48, "fov": 56, "cam_yaw": 64, "cam_pitch": 72,
              "cam_roll": 80, "disk_enabled": 88, "disk_inner": 96, "disk_outer": 104,
              "disk_temperature": 112, "disk_volumetric": 120, "disk_alpha": 128,
              "disk_turbulence": 136, "disk_seed": 144, "disk_noise_scale": 152,
              "disk_noise_octaves": 160, "mass_solar": 168, "eddington_fraction": 176,
              "background_type": 184, "background_texture_path": 192,
              "integrator_tolerance": 200, "integrator_max_steps": 208,
              "samples_per_pixel": 212, "thread_count": 216, "backend": 220,
              "num_frequency_bins": 224, "frequency_bins_hz": 232}
    assert offs == expect


def test_no_gpu_fails_loudly():
    import grrt_amd
    if grrt_amd.gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="grrt_create failed"):
        grrt_amd.Context(grrt_amd.make_params(width=8, height=8))
    # the GPU table build refuses too; the host build needs no GPU
    p = grrt_amd.make_params(metric=1, spin=0.9, volumetric=True, disk_outer=20.0)
    with pytest.raises(RuntimeError, match="volumetric table build failed"):
        grrt_amd.volumetric_tables(p, device=True)
    assert grrt_amd.volumetric_tables(p)["n_nu"] == 20


def test_tonemap_host_only():
    """grrt_tonemap is host post-processing (src/tonemapper.cpp:7-71):
    log-average exposure, extended Reinhard, gamma 2.2."""
    import numpy as np

    import grrt_amd
    rng = np.random.default_rng(0)
    fb = np.zeros((8, 8, 4), np.float32)
    fb[..., :3] = rng.lognormal(size=(8, 8, 3)).astype(np.float32)
    fb[0, 0, :3] = 0.0
    out = grrt_amd.tonemap(fb)
    assert np.all(out[..., :3] >= 0) and np.all(out[..., :3] <= 1.0)
    assert np.all(out[0, 0, :3] == 0)
    # chrominance is kept: undoing gamma gives one scale factor per pixel
    lin = out[..., :3].astype(np.float64) ** 2.2
    ok = np.all(lin < 0.999, axis=-1) & np.all(fb[..., :3] > 0, axis=-1)
    ratio = lin[ok] / fb[..., :3][ok]
    np.testing.assert_allclose(ratio, ratio[:, :1].repeat(3, 1), rtol=1e-4)
    # the mapped luminance is monotone in the input luminance
    lum = 0.2126 * fb[..., 0] + 0.7152 * fb[..., 1] + 0.0722 * fb[..., 2]
    lum_o = (lin * np.array([0.2126, 0.7152, 0.0722])).sum(-1)
    order = np.argsort(lum[ok])
    assert np.all(np.diff(lum_o[ok][order]) >= -1e-6)
    black = np.zeros((4, 4, 4), np.float32)
    assert np.array_equal(grrt_amd.tonemap(black), black)


def test_tonemap_closed_forms():
    """Cases the curve fixes exactly: a uniform grey frame has its white
    point at its own luminance, so every channel maps to 1.0; scaling a frame
    by k scales the log-average exposure by 1/k, so the output is unchanged
    while k L stays far above the 1e-6 offset of the log average."""
    import numpy as np

    import grrt_amd
    grey = np.zeros((5, 7, 4), np.float32)
    grey[..., :3] = 0.37
    grey[..., 3] = 1.0
    out = grrt_amd.tonemap(grey)
    np.testing.assert_allclose(out[..., :3], 1.0, rtol=0, atol=2e-7)
    assert np.array_equal(out[..., 3], grey[..., 3])
    rng = np.random.default_rng(5)
    fb = np.zeros((16, 16, 4), np.float32)
    fb[..., :3] = rng.lognormal(0.0, 1.0, (16, 16, 3)).astype(np.float32)
    a = grrt_amd.tonemap(fb)
    b = grrt_amd.tonemap(fb * np.float32(64.0))
    np.testing.assert_allclose(a[..., :3], b[..., :3], rtol=0, atol=2e-6)


C_CALLER = r"""
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include "grrt/api.h"
#include "grrt_ext.h"

_Static_assert(sizeof(GRRTParams) == 240, "GRRTParams is 240 bytes on LP64");
_Static_assert(offsetof(GRRTParams, background_texture_path) == 192, "layout");
_Static_assert(offsetof(GRRTParams, frequency_bins_hz) == 232, "layout");

int main(void) {
    GRRTParams p;
    memset(&p, 0, sizeof p);
    p.width = 8;
    p.height = 8;
    p.backend = GRRT_BACKEND_CUDA;
    if (!grrt_cuda_available()) {
        GRRTContext* ctx = grrt_create(&p);
        if (ctx != NULL) return 2;
        const char* e = grrt_last_error();
        if (e == NULL || e[0] == 0) return 3;
    }
    float fb[4 * 4] = {1.0f, 0.5f, 0.25f, 1.0f, 2.0f, 1.0f, 0.5f, 1.0f,
                       0.0f, 0.0f, 0.0f, 1.0f, 4.0f, 4.0f, 4.0f, 1.0f};
    grrt_tonemap(fb, 2, 2);
    for (int i = 0; i < 16; ++i)
        if (!(fb[i] >= 0.0f && fb[i] <= 1.0f)) return 4;
    printf("ok\n");
    return 0;
}
"""


def test_plain_c_caller_compiles_links_and_runs(tmp_path):
    """A C11 program (as a C host such as the reference's CLI or a cgo
    binding would be) includes both headers with -Wall -Werror, checks the
    GRRTParams layout at compile time, links against libgrrt.so and runs:
    without a GPU grrt_create returns NULL with a reason; the host tone map
    works either way."""
    import shutil
    import subprocess
    import grrt_amd
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(grrt_amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", libdir, "-lgrrt", f"-Wl,-rpath,{libdir}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)
