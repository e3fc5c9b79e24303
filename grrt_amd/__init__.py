"""Python binding of libgrrt.so, the MI355X GRRT backend.

The binding mirrors the reference's C ABI (include/grrt/api.h) one to one
through ctypes, the way the reference's planned Blender binding would have
(gr-raytracer-blender-spec_1.md:566-600), plus the ``grrt_ext_*`` entry
points for sampling control, per-ray records and counters.  There is no
CPU fallback: if the library is not built, or no GPU is present,
``Context(...)`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgrrt.so")

METRIC_SCHWARZSCHILD, METRIC_KERR = 0, 1
BACKEND_CPU, BACKEND_GPU = 0, 1
BG_BLACK, BG_STARS, BG_TEXTURE = 0, 1, 2
SAMPLING_HASH, SAMPLING_CENTER, SAMPLING_SEEDED = 0, 1, 2
CLASS_CAPTURED, CLASS_ESCAPED, CLASS_DISK, CLASS_MAXSTEPS = 0, 1, 2, 3


class GRRTParams(C.Structure):
    """Field-for-field copy of GRRTParams (include/grrt/types.h)."""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("metric_type", C.c_int), ("mass", C.c_double), ("spin", C.c_double),
        ("observer_r", C.c_double), ("observer_theta", C.c_double),
        ("observer_phi", C.c_double), ("fov", C.c_double),
        ("cam_yaw", C.c_double), ("cam_pitch", C.c_double), ("cam_roll", C.c_double),
        ("disk_enabled", C.c_int), ("disk_inner", C.c_double), ("disk_outer", C.c_double),
        ("disk_temperature", C.c_double),
        ("disk_volumetric", C.c_int), ("disk_alpha", C.c_double),
        ("disk_turbulence", C.c_double), ("disk_seed", C.c_int),
        ("disk_noise_scale", C.c_double), ("disk_noise_octaves", C.c_int),
        ("mass_solar", C.c_double), ("eddington_fraction", C.c_double),
        ("background_type", C.c_int), ("background_texture_path", C.c_char_p),
        ("integrator_tolerance", C.c_double), ("integrator_max_steps", C.c_int),
        ("samples_per_pixel", C.c_int), ("thread_count", C.c_int), ("backend", C.c_int),
        ("num_frequency_bins", C.c_int), ("frequency_bins_hz", C.POINTER(C.c_double)),
    ]


assert C.sizeof(GRRTParams) == 240


class GRRTRayRecord(C.Structure):
    _fields_ = [
        ("color", C.c_double * 3), ("x_final", C.c_double * 4), ("p_final", C.c_double * 4),
        ("r_hit", C.c_double), ("g_hit", C.c_double), ("h_max", C.c_double),
        ("h_final", C.c_double),
        ("term", C.c_int32), ("cls", C.c_int32), ("steps", C.c_int32),
        ("attempts", C.c_int32), ("n_cross", C.c_int32), ("pad", C.c_int32),
    ]


RECORD_DTYPE = np.dtype([
    ("color", "<f8", 3), ("x_final", "<f8", 4), ("p_final", "<f8", 4),
    ("r_hit", "<f8"), ("g_hit", "<f8"), ("h_max", "<f8"), ("h_final", "<f8"),
    ("term", "<i4"), ("cls", "<i4"), ("steps", "<i4"), ("attempts", "<i4"),
    ("n_cross", "<i4"), ("pad", "<i4"),
])
assert RECORD_DTYPE.itemsize == C.sizeof(GRRTRayRecord) == 144


class GRRTStats(C.Structure):
    _fields_ = [
        ("trace_ms", C.c_double), ("finalize_ms", C.c_double), ("total_ms", C.c_double),
        ("rays", C.c_int64), ("steps", C.c_int64), ("attempts", C.c_int64),
        ("deriv_evals", C.c_int64), ("algo_flops", C.c_double),
        ("cancelled", C.c_int32), ("pad", C.c_int32),
        ("march_steps", C.c_int64), ("copy_ms", C.c_double),
    ]


class GRRTView(C.Structure):
    _fields_ = [
        ("spin", C.c_double), ("observer_r", C.c_double), ("observer_theta", C.c_double),
        ("observer_phi", C.c_double), ("fov", C.c_double), ("width", C.c_int), ("height", C.c_int),
    ]


def views_array(views) -> "C.Array":
    """[(spin, r, theta, phi, fov, width, height) or dict] -> GRRTView array."""
    arr = (GRRTView * len(views))()
    for k, v in enumerate(views):
        if isinstance(v, dict):
            for f, _ in GRRTView._fields_:
                if f in v:
                    setattr(arr[k], f, v[f])
        else:
            for (f, _), x in zip(GRRTView._fields_, v):
                setattr(arr[k], f, x)
    return arr


class GRRTTrainConfig(C.Structure):
    _fields_ = [
        ("steps", C.c_int), ("batch", C.c_int), ("lr_max", C.c_double), ("lr_min", C.c_double),
        ("warmup", C.c_int), ("pad", C.c_int), ("w_r", C.c_double), ("w_g", C.c_double),
        ("w_dir", C.c_double), ("seed", C.c_uint64),
    ]


class GRRTTrainStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("rays", C.c_int64), ("loss", C.c_double), ("ce", C.c_double),
        ("accuracy", C.c_double), ("rmse_lnr", C.c_double), ("rmse_lng", C.c_double),
        ("rmse_dir", C.c_double), ("ms", C.c_double),
    ]


class GRRTSurrogateStats(C.Structure):
    _fields_ = [
        ("infer_ms", C.c_double), ("trace_ms", C.c_double), ("gpu_ms", C.c_double),
        ("total_ms", C.c_double), ("rays", C.c_int64), ("fallback_rays", C.c_int64),
        ("judged_rays", C.c_int64), ("trace_attempts", C.c_int64),
        ("judge_class_agreement", C.c_double), ("judge_median_rel_err", C.c_double),
    ]


def _as_dict(st: C.Structure) -> dict:
    return {f: getattr(st, f) for f, _ in st._fields_ if f != "pad"}


SURROGATE_PARAMS = 201225

PROGRESS_FN = C.CFUNCTYPE(None, C.c_float, C.c_void_p)

# Every symbol the C headers declare: (name, restype, argtypes).
_P = C.c_void_p
API = [
    ("grrt_create", _P, [C.POINTER(GRRTParams)]),
    ("grrt_destroy", None, [_P]),
    ("grrt_update_params", None, [_P, C.POINTER(GRRTParams)]),
    ("grrt_render", C.c_int, [_P, C.POINTER(C.c_float)]),
    ("grrt_render_cb", C.c_int, [_P, C.POINTER(C.c_float), PROGRESS_FN, _P]),
    ("grrt_render_tile", C.c_int, [_P, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.c_int]),
    ("grrt_cancel", None, [_P]),
    ("grrt_progress", C.c_float, [_P]),
    ("grrt_error", C.c_char_p, [_P]),
    ("grrt_cuda_available", C.c_int, []),
    ("grrt_last_error", C.c_char_p, []),
    ("grrt_tonemap", None, [C.POINTER(C.c_float), C.c_int, C.c_int]),
    ("grrt_debug_pixel", None, [_P, C.c_int, C.c_int]),
    ("grrt_set_frequency_bins", None, [_P, C.POINTER(C.c_double), C.c_int]),
    ("grrt_render_spectral", C.c_int, [_P, C.POINTER(C.c_double), C.c_int, C.c_int]),
    ("grrt_render_spectral_cb", C.c_int,
     [_P, C.POINTER(C.c_double), C.c_int, C.c_int, PROGRESS_FN, _P]),
    ("grrt_render_spectral_to_fits_cb", C.c_int,
     [_P, C.c_char_p, C.c_int, C.c_int, PROGRESS_FN, _P]),
    ("grrt_ext_set_sampling", C.c_int, [_P, C.c_int, C.c_int, C.c_uint64]),
    ("grrt_ext_set_records", C.c_int, [_P, C.c_int]),
    ("grrt_ext_num_samples", C.c_int64, [_P]),
    ("grrt_ext_get_records", C.c_int, [_P, _P, C.c_int64]),
    ("grrt_ext_get_classes", C.c_int, [_P, _P, C.c_int64]),
    ("grrt_ext_last_stats", C.c_int, [_P, C.POINTER(GRRTStats)]),
    ("grrt_ext_render_device", C.c_int, [_P, _P]),
    ("grrt_ext_render_device_noalloc", C.c_int, [_P, _P]),
    ("grrt_ext_set_tuning", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("grrt_ext_render_batch", C.c_int, [_P, C.POINTER(GRRTView), C.c_int, _P, C.c_int]),
    ("grrt_ext_autotune", C.c_int, [_P, C.POINTER(GRRTView), C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]),
    ("grrt_ext_set_precision", C.c_int, [_P, C.c_int]),
    ("grrt_ext_render_spectral_to_fits_f32", C.c_int, [_P, C.c_char_p, C.c_int, C.c_int, PROGRESS_FN, _P]),
    ("grrt_ext_write_fits_cube", C.c_int, [C.c_char_p, _P, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_double), C.POINTER(GRRTParams)]),
    ("grrt_ext_write_hdr", C.c_int, [C.c_char_p, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int]),
    ("grrt_ext_read_texture", C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float),
                                        C.c_int64]),
    ("grrt_ext_render_hdr", C.c_int, [_P, C.c_char_p, C.c_char_p, C.POINTER(C.c_float), PROGRESS_FN, C.c_void_p]),
    ("grrt_ext_surrogate_init", C.c_int, [_P, C.c_uint64]),
    ("grrt_ext_surrogate_generate", C.c_int, [_P, C.c_int64, C.c_uint64, C.c_int]),
    ("grrt_ext_surrogate_train", C.c_int, [_P, C.POINTER(GRRTTrainConfig), C.POINTER(GRRTTrainStats)]),
    ("grrt_ext_surrogate_evaluate", C.c_int, [_P, C.POINTER(GRRTTrainStats)]),
    ("grrt_ext_surrogate_params", C.c_int, [_P, _P, C.c_int64, C.c_int]),
    ("grrt_ext_surrogate_forward", C.c_int, [_P, _P, C.c_int64, _P]),
    ("grrt_ext_surrogate_gradients", C.c_int, [_P, _P, _P, C.c_int64, _P, _P, C.POINTER(C.c_double)]),
    ("grrt_ext_surrogate_dataset", C.c_int, [_P, C.c_int, _P, _P, C.c_int64]),
    ("grrt_ext_surrogate_set_critical_gate", C.c_int, [_P, C.c_double]),
    ("grrt_ext_render_surrogate", C.c_int, [_P, _P, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.POINTER(GRRTSurrogateStats)]),
    ("grrt_ext_surrogate_infer_raw", C.c_int, [_P, _P, C.c_int64, _P]),
    ("grrt_ext_surrogate_adam_step", C.c_int, [_P, _P, _P, _P, _P, C.c_int64, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_int64]),
    ("grrt_ext_surrogate_lists", C.c_int, [_P, _P, C.c_int64, _P, C.c_int64]),
    ("grrt_ext_get_samples", C.c_int, [_P, _P, C.c_int64]),
    ("grrt_ext_tonemap_device", C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    ("grrt_ext_get_frame", C.c_int, [_P, _P]),
    ("grrt_ext_set_integrator", C.c_int, [_P, C.c_int]),
    ("grrt_ext_render_spectral_f32", C.c_int, [_P, C.POINTER(C.c_float), C.c_int, C.c_int]),
    ("grrt_ext_render_spectral_tile", C.c_int, [_P, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int]),
    ("grrt_ext_set_vol_tuning", C.c_int, [_P, C.c_int]),
    ("grrt_ext_vol_build", _P, [C.POINTER(GRRTParams)]),
    ("grrt_ext_vol_build_device", _P, [C.POINTER(GRRTParams)]),
    ("grrt_ext_vol_build_ex", _P, [C.POINTER(GRRTParams), C.c_int]),
    ("grrt_ext_vol_scalars", C.c_int, [_P, C.POINTER(C.c_double), C.c_int]),
    ("grrt_ext_vol_table", C.c_int64, [_P, C.c_char_p, C.POINTER(C.c_double), C.c_int64]),
    ("grrt_ext_vol_free", None, [_P]),
    ("grrt_ext_resolve_params", C.c_int, [C.POINTER(GRRTParams), C.POINTER(C.c_double), C.c_int]),
    ("grrt_ext_saha", C.c_int, [C.c_double, C.c_double, C.POINTER(C.c_double)]),
    ("grrt_ext_opacity", C.c_int, [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]),
]

VOL_SCALARS = ["r_isco", "r_horizon", "r_min", "r_outer", "taper_width", "E_isco", "L_isco",
               "rho_scale", "noise_scale", "turbulence", "octaves", "n_r", "n_z", "n_nu",
               "log_nu_min", "log_nu_max", "M", "a", "t_peak", "build_ms_opacity", "build_ms_total",
               "build_ms_columns", "build_ms_columns_kernel"]
VOL_TABLES = ["H", "z_max", "rho_mid", "T_eff", "rho_profile", "T_profile", "kappa_abs",
              "kappa_es", "kappa_ross", "mu", "perm"]

_lib: C.CDLL | None = None


def build(jobs: int = 8) -> None:
    """Compile libgrrt.so for gfx950 with the committed Makefile."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(_HERE, "csrc")], check=True)


def lib() -> C.CDLL:
    """Load libgrrt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run grrt_amd.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in API:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def gpu_available() -> bool:
    return bool(lib().grrt_cuda_available())


def make_params(width=256, height=256, metric=METRIC_KERR, spin=0.998, mass=1.0,
                observer_r=50.0, observer_theta=1.3962634, observer_phi=0.0,
                fov=1.5707963, disk=True, disk_inner=0.0, disk_outer=20.0,
                disk_temperature=0.0, stars=False, tolerance=1e-8, max_steps=10000,
                spp=1, volumetric=False, alpha=0.1, turbulence=0.4, seed=42,
                noise_scale=0.0, octaves=2, freqs=None, texture=None) -> GRRTParams:
    p = GRRTParams()
    p.width, p.height = width, height
    p.metric_type = metric
    p.mass, p.spin = mass, spin
    p.observer_r, p.observer_theta, p.observer_phi, p.fov = observer_r, observer_theta, observer_phi, fov
    p.disk_enabled = 1 if disk else 0
    p.disk_inner, p.disk_outer, p.disk_temperature = disk_inner, disk_outer, disk_temperature
    p.disk_volumetric = 1 if volumetric else 0
    p.background_type = BG_STARS if stars else BG_BLACK
    if texture is not None:
        p.background_type = BG_TEXTURE
        p._texture_keepalive = texture.encode()
        p.background_texture_path = p._texture_keepalive
    p.integrator_tolerance, p.integrator_max_steps = tolerance, max_steps
    p.samples_per_pixel = spp
    p.backend = BACKEND_GPU
    p.disk_alpha, p.disk_turbulence, p.disk_seed = alpha, turbulence, seed
    p.disk_noise_scale, p.disk_noise_octaves = noise_scale, octaves
    if freqs is not None:
        arr = (C.c_double * len(freqs))(*freqs)
        p._freqs_keepalive = arr
        p.num_frequency_bins = len(freqs)
        p.frequency_bins_hz = C.cast(arr, C.POINTER(C.c_double))
    return p


def volumetric_tables(params: GRRTParams, device: bool = False, device_columns: bool | None = None) -> dict:
    """Volumetric disk set-up: scalars and tables.  device=False builds on
    host threads (no GPU needed); device=True computes the opacity tables and
    the vertical columns on the GPU, as grrt_create does (device_columns=False:
    the columns on host threads over the GPU's opacity tables)."""
    L = lib()
    if device and device_columns is False:
        h = L.grrt_ext_vol_build_ex(C.byref(params), 1)
    else:
        fn = L.grrt_ext_vol_build_device if device else L.grrt_ext_vol_build
        h = fn(C.byref(params))
    if not h:
        err = L.grrt_last_error()
        raise RuntimeError("volumetric table build failed: " + (err.decode() if err else "unknown"))
    try:
        sc = np.zeros(len(VOL_SCALARS))
        L.grrt_ext_vol_scalars(h, sc.ctypes.data_as(C.POINTER(C.c_double)), len(sc))
        out = dict(zip(VOL_SCALARS, sc.tolist()))
        for name in VOL_TABLES:
            n = L.grrt_ext_vol_table(h, name.encode(), None, 0)
            arr = np.zeros(n)
            L.grrt_ext_vol_table(h, name.encode(), arr.ctypes.data_as(C.POINTER(C.c_double)), n)
            out[name] = arr
        out["perm"] = out["perm"].astype(np.int32)
        return out
    finally:
        L.grrt_ext_vol_free(h)


RESOLVED = ["width", "height", "M", "a", "r_obs", "theta_obs", "phi_obs", "fov", "max_steps", "tol",
            "disk", "volumetric", "r_in", "r_out", "t_disk", "stars", "spp", "r_horizon", "r_isco",
            "vol_M", "vol_spin", "vol_r_outer", "vol_t_peak", "vol_nu_min", "vol_nu_max"]


def resolve_params(params: GRRTParams) -> dict:
    """The scene grrt_create would build from params (no GPU needed)."""
    out = np.zeros(len(RESOLVED))
    n = lib().grrt_ext_resolve_params(C.byref(params), out.ctypes.data_as(C.POINTER(C.c_double)), len(out))
    if n != len(RESOLVED):
        raise RuntimeError("grrt_ext_resolve_params failed")
    return dict(zip(RESOLVED, out.tolist()))


def saha(rho, T) -> dict:
    out = np.zeros(5)
    lib().grrt_ext_saha(rho, T, out.ctypes.data_as(C.POINTER(C.c_double)))
    return dict(zip(["n_e", "mu", "n_ion_eff", "n_HI", "n_Hminus"], out.tolist()))


def opacity(nu, rho, T) -> dict:
    out = np.zeros(6)
    lib().grrt_ext_opacity(nu, rho, T, out.ctypes.data_as(C.POINTER(C.c_double)))
    return dict(zip(["kappa_abs", "kappa_es", "alpha_ff", "alpha_hminus", "alpha_bf", "B_nu"],
                    out.tolist()))


class Context:
    """Owns one GRRTContext (grrt_create ... grrt_destroy)."""

    def __init__(self, params: GRRTParams):
        self._L = lib()
        self.params = params
        self._h = self._L.grrt_create(C.byref(params))
        if not self._h:
            err = self._L.grrt_last_error()
            raise RuntimeError("grrt_create failed: " + (err.decode() if err else "unknown"))

    def close(self):
        if self._h:
            self._L.grrt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            err = self._L.grrt_error(self._h)
            raise RuntimeError(f"{what} failed: " + (err.decode() if err else "unknown"))
        return rc

    @property
    def error(self):
        e = self._L.grrt_error(self._h)
        return e.decode() if e else None

    def set_sampling(self, mode: int, spp: int = 1, seed: int = 0):
        self._check(self._L.grrt_ext_set_sampling(self._h, mode, spp, seed), "set_sampling")

    def set_records(self, enable: bool = True):
        self._check(self._L.grrt_ext_set_records(self._h, 1 if enable else 0), "set_records")

    def set_tuning(self, block=0, blocks_per_cu=0, refill_min=0, chunk=0, tile=-1):
        self._check(self._L.grrt_ext_set_tuning(self._h, block, blocks_per_cu, refill_min, chunk, tile),
                    "set_tuning")

    def render_batch(self, views, device: bool = False) -> list[np.ndarray] | None:
        """Render several views in one persistent launch (grrt_ext_render_batch);
        returns one (height, width, 4) float32 frame per view, or None with
        device=True (frames stay in the context's device memory)."""
        arr = views_array(views)
        if device:
            self._check(self._L.grrt_ext_render_batch(self._h, arr, len(arr), None, 1), "grrt_ext_render_batch")
            return None
        npx = sum(v.width * v.height for v in arr)
        out = np.zeros(4 * npx, np.float32)
        self._check(self._L.grrt_ext_render_batch(self._h, arr, len(arr), out.ctypes.data, 0),
                    "grrt_ext_render_batch")
        frames, o = [], 0
        for v in arr:
            n = v.width * v.height * 4
            frames.append(out[o:o + n].reshape(v.height, v.width, 4))
            o += n
        return frames

    def render_spectral_to_fits(self, path: str, f32: bool = False, progress=None) -> int:
        """Stream the spectral cube to a FITS file (BITPIX -64, or -32 with
        f32); returns 0, or 1 if cancelled."""
        cb = PROGRESS_FN(lambda f, _ud: progress(f)) if progress else PROGRESS_FN(0)
        fn = self._L.grrt_ext_render_spectral_to_fits_f32 if f32 else self._L.grrt_render_spectral_to_fits_cb
        return self._check(fn(self._h, path.encode(), self.params.width, self.params.height, cb, None),
                           "render_spectral_to_fits")

    def render_hdr(self, path_linear: str, path_scaled: str | None = None, progress=None):
        """grrt_ext_render_hdr: render with the rows streamed to the linear
        Radiance file (and the exposure-scaled one after it); returns
        (rc, frame) with rc 0, or 1 if cancelled."""
        fb = np.zeros((self.params.height, self.params.width, 4), np.float32)
        cb = PROGRESS_FN(lambda f, _ud: progress(f)) if progress else PROGRESS_FN(0)
        rc = self._check(self._L.grrt_ext_render_hdr(self._h, path_linear.encode(),
                                                      path_scaled.encode() if path_scaled else None,
                                                      fb.ctypes.data_as(C.POINTER(C.c_float)), cb, None),
                         "render_hdr")
        return rc, fb

    # ---- learned surrogate ----
    def surrogate_init(self, seed: int = 0):
        self._check(self._L.grrt_ext_surrogate_init(self._h, seed), "surrogate_init")

    def surrogate_generate(self, n_rays: int, seed: int = 0, validation: bool = False):
        self._check(self._L.grrt_ext_surrogate_generate(self._h, n_rays, seed, 1 if validation else 0),
                    "surrogate_generate")

    def surrogate_train(self, steps, batch=4096, lr_max=1e-3, lr_min=1e-5, warmup=100, w_r=10.0, w_g=10.0,
                        w_dir=1.0, seed=0) -> dict:
        cfg = GRRTTrainConfig(steps=steps, batch=batch, lr_max=lr_max, lr_min=lr_min, warmup=warmup,
                              w_r=w_r, w_g=w_g, w_dir=w_dir, seed=seed)
        st = GRRTTrainStats()
        self._check(self._L.grrt_ext_surrogate_train(self._h, C.byref(cfg), C.byref(st)), "surrogate_train")
        return _as_dict(st)

    def surrogate_evaluate(self) -> dict:
        st = GRRTTrainStats()
        self._check(self._L.grrt_ext_surrogate_evaluate(self._h, C.byref(st)), "surrogate_evaluate")
        return _as_dict(st)

    def surrogate_params(self, values: np.ndarray | None = None) -> np.ndarray:
        if values is None:
            out = np.zeros(SURROGATE_PARAMS, np.float32)
            self._check(self._L.grrt_ext_surrogate_params(self._h, out.ctypes.data, out.size, 0), "params")
            return out
        v = np.ascontiguousarray(values, np.float32)
        self._check(self._L.grrt_ext_surrogate_params(self._h, v.ctypes.data, v.size, 1), "params")
        return v

    def surrogate_forward(self, x: np.ndarray) -> np.ndarray:
        xx = np.ascontiguousarray(x, np.float32)
        y = np.zeros((xx.shape[0], 9), np.float32)
        self._check(self._L.grrt_ext_surrogate_forward(self._h, xx.ctypes.data, xx.shape[0], y.ctypes.data),
                    "surrogate_forward")
        return y

    def surrogate_gradients(self, x, t, w=(1.0, 1.0, 1.0)):
        xx, tt = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(t, np.float32)
        ww = np.ascontiguousarray(w, np.float32)
        g = np.zeros(SURROGATE_PARAMS, np.float32)
        loss = C.c_double()
        self._check(self._L.grrt_ext_surrogate_gradients(self._h, xx.ctypes.data, tt.ctypes.data, xx.shape[0],
                                                          ww.ctypes.data, g.ctypes.data, C.byref(loss)),
                    "surrogate_gradients")
        return loss.value, g

    def surrogate_dataset(self, n: int, validation: bool = False):
        x = np.zeros((n, 5), np.float32)
        t = np.zeros((n, 8), np.float32)
        self._check(self._L.grrt_ext_surrogate_dataset(self._h, 1 if validation else 0, x.ctypes.data,
                                                        t.ctypes.data, n), "surrogate_dataset")
        return x, t

    def set_critical_gate(self, score_max: float) -> None:
        """grrt_ext_surrogate_set_critical_gate: rays with |critical score|
        below score_max fall back to the integrator (0: off)."""
        self._check(self._L.grrt_ext_surrogate_set_critical_gate(self._h, score_max), "set_critical_gate")

    def render_surrogate(self, entropy_max=0.1, judge_fraction=0.02, lng_tol=0.0, d_ptr: int | None = None,
                         device: bool = False):
        """Surrogate frame with FP64 fallback and judge; returns (frame, stats
        dict).  d_ptr: render into that device buffer; device=True: keep the
        frame in the context's device memory (frame None)."""
        st = GRRTSurrogateStats()
        if device:
            self._check(self._L.grrt_ext_render_surrogate(self._h, None, 1, entropy_max, lng_tol,
                                                           judge_fraction, C.byref(st)), "render_surrogate")
            return None, _as_dict(st)
        if d_ptr is not None:
            self._check(self._L.grrt_ext_render_surrogate(self._h, d_ptr, 1, entropy_max, lng_tol,
                                                           judge_fraction, C.byref(st)), "render_surrogate")
            return None, _as_dict(st)
        fb = np.zeros((self.params.height, self.params.width, 4), np.float32)
        self._check(self._L.grrt_ext_render_surrogate(self._h, fb.ctypes.data, 0, entropy_max, lng_tol,
                                                       judge_fraction, C.byref(st)), "render_surrogate")
        return fb, _as_dict(st)

    def surrogate_infer_raw(self, x: np.ndarray | None = None) -> np.ndarray:
        """The surrogate frame's network pass, raw outputs [n][9]: on inputs x
        [n][5], or on the frame's own samples (x None)."""
        if x is None:
            n, xp = self.num_samples(), None
        else:
            xx = np.ascontiguousarray(x, np.float32)
            n, xp = xx.shape[0], xx.ctypes.data
        y = np.zeros((n, 9), np.float32)
        self._check(self._L.grrt_ext_surrogate_infer_raw(self._h, xp, n, y.ctypes.data), "surrogate_infer_raw")
        return y

    def surrogate_adam_step(self, w, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
        """One step of the training loop's Adam kernel, in place on float32
        arrays w, m, v."""
        for a in (w, m, v):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        gg = np.ascontiguousarray(g, np.float32)
        self._check(self._L.grrt_ext_surrogate_adam_step(self._h, w.ctypes.data, gg.ctypes.data, m.ctypes.data,
                                                          v.ctypes.data, w.size, lr, b1, b2, eps, t), "adam_step")

    def surrogate_lists(self, n_fallback: int, n_judged: int):
        fb = np.zeros(n_fallback, np.int64)
        jd = np.zeros(n_judged, np.int64)
        self._check(self._L.grrt_ext_surrogate_lists(self._h, fb.ctypes.data, n_fallback, jd.ctypes.data,
                                                      n_judged), "surrogate_lists")
        return fb, jd

    def samples(self, n: int | None = None) -> np.ndarray:
        """Per-sample linear RGB of the last frame, [n][3] float64."""
        n = self.num_samples() if n is None else n
        out = np.zeros((n, 3), np.float64)
        self._check(self._L.grrt_ext_get_samples(self._h, out.ctypes.data, n), "get_samples")
        return out

    def autotune(self, views=None) -> tuple[dict, float]:
        """Time launch shapes on the frame (views None) or the batch and keep
        the fastest; returns ({tile, blocks_per_cu, refill_min, chunk}, ms)."""
        best = (C.c_int * 4)()
        ms = C.c_double()
        if views is None:
            rc = self._L.grrt_ext_autotune(self._h, None, 0, best, C.byref(ms))
        else:
            arr = views_array(views)
            rc = self._L.grrt_ext_autotune(self._h, arr, len(arr), best, C.byref(ms))
        self._check(rc, "grrt_ext_autotune")
        return dict(zip(("tile", "blocks_per_cu", "refill_min", "chunk"), list(best))), ms.value

    def render(self, progress=None, out: np.ndarray | None = None) -> np.ndarray:
        """grrt_render / grrt_render_cb into a host frame (out, if given, is
        reused: a (height, width, 4) C-contiguous float32 array)."""
        if out is None:
            fb = np.zeros((self.params.height, self.params.width, 4), np.float32)
        else:
            assert out.dtype == np.float32 and out.flags.c_contiguous
            assert out.shape == (self.params.height, self.params.width, 4)
            fb = out
        ptr = fb.ctypes.data_as(C.POINTER(C.c_float))
        if progress is None:
            rc = self._L.grrt_render(self._h, ptr)
        else:
            cb = PROGRESS_FN(lambda f, _ud: progress(f))
            rc = self._L.grrt_render_cb(self._h, ptr, cb, None)
        self._check(rc, "grrt_render")
        self.cancelled = rc == 1
        return fb

    def render_tile(self, x, y, w, h) -> np.ndarray:
        buf = np.zeros((h, w, 4), np.float32)
        self._check(self._L.grrt_render_tile(self._h, buf.ctypes.data_as(C.POINTER(C.c_float)),
                                             x, y, w, h), "grrt_render_tile")
        return buf

    def render_device(self, d_ptr: int | None = None) -> int:
        return self._check(self._L.grrt_ext_render_device(self._h, d_ptr), "grrt_ext_render_device")

    def render_device_noalloc(self, d_ptr: int) -> int:
        """grrt_ext_render_device_noalloc into the device buffer d_ptr."""
        return self._check(self._L.grrt_ext_render_device_noalloc(self._h, d_ptr), "grrt_ext_render_device_noalloc")

    def tonemap_device(self, d_ptr: int | None = None) -> float:
        """Tone-map the context's device frame (or d_ptr) in place; returns
        the kernels' GPU time in ms."""
        ms = C.c_double()
        self._check(self._L.grrt_ext_tonemap_device(self._h, d_ptr, C.byref(ms)), "grrt_ext_tonemap_device")
        return ms.value

    def get_frame(self) -> np.ndarray:
        """The context's device frame as a (height, width, 4) float32 array."""
        out = np.zeros((self.params.height, self.params.width, 4), np.float32)
        self._check(self._L.grrt_ext_get_frame(self._h, out.ctypes.data), "grrt_ext_get_frame")
        return out

    def num_samples(self) -> int:
        return int(self._L.grrt_ext_num_samples(self._h))

    def classes(self, n: int | None = None) -> np.ndarray:
        n = self.num_samples() if n is None else n
        out = np.zeros(n, np.uint8)
        self._check(self._L.grrt_ext_get_classes(self._h, out.ctypes.data, n), "get_classes")
        return out

    def records(self, n: int | None = None) -> np.ndarray:
        n = self.num_samples() if n is None else n
        out = np.zeros(n, RECORD_DTYPE)
        self._check(self._L.grrt_ext_get_records(self._h, out.ctypes.data, n), "get_records")
        return out

    def stats(self) -> GRRTStats:
        s = GRRTStats()
        self._check(self._L.grrt_ext_last_stats(self._h, C.byref(s)), "last_stats")
        return s

    def cancel(self):
        self._L.grrt_cancel(self._h)

    def progress(self) -> float:
        return float(self._L.grrt_progress(self._h))

    def update_params(self, params: GRRTParams):
        self.params = params
        self._L.grrt_update_params(self._h, C.byref(params))
        if self.error:
            raise RuntimeError("grrt_update_params failed: " + self.error)

    def set_frequency_bins(self, freqs):
        arr = np.ascontiguousarray(freqs, np.float64)
        self._freqs = arr
        self._L.grrt_set_frequency_bins(self._h, arr.ctypes.data_as(C.POINTER(C.c_double)), len(arr))

    def set_precision(self, mode):
        """0 / False: all FP64; 1 / True: FP32 transfer; 2: FP32 transfer and
        FP32 far-field geodesic steps (volumetric path)."""
        self._check(self._L.grrt_ext_set_precision(self._h, int(mode)), "set_precision")

    def set_vol_tuning(self, waves_per_simd: int):
        """Volumetric tracer: 1 or 2 persistent wavefronts per SIMD."""
        self._check(self._L.grrt_ext_set_vol_tuning(self._h, int(waves_per_simd)), "set_vol_tuning")

    def set_integrator(self, mode: int):
        """0: adaptive Dormand-Prince (default); 1: step-doubling RK4."""
        self._check(self._L.grrt_ext_set_integrator(self._h, int(mode)), "set_integrator")

    def render_spectral_f32(self, width=None, height=None) -> np.ndarray:
        w = self.params.width if width is None else width
        h = self.params.height if height is None else height
        n = len(getattr(self, "_freqs", []))
        buf = np.zeros((h, w, max(n, 1)), np.float32)
        self._check(self._L.grrt_ext_render_spectral_f32(
            self._h, buf.ctypes.data_as(C.POINTER(C.c_float)), w, h), "render_spectral_f32")
        return buf

    def render_spectral_tile(self, x, y, w, h) -> np.ndarray:
        n = len(getattr(self, "_freqs", []))
        buf = np.zeros((h, w, max(n, 1)), np.float64)
        self._check(self._L.grrt_ext_render_spectral_tile(self._h, buf.ctypes.data_as(C.POINTER(C.c_double)),
                                                           x, y, w, h), "render_spectral_tile")
        return buf

    def render_spectral(self, width=None, height=None) -> np.ndarray:
        w = self.params.width if width is None else width
        h = self.params.height if height is None else height
        n = len(getattr(self, "_freqs", []))
        buf = np.zeros((h, w, max(n, 1)), np.float64)
        self._check(self._L.grrt_render_spectral(self._h, buf.ctypes.data_as(C.POINTER(C.c_double)),
                                                 w, h), "grrt_render_spectral")
        return buf


def tonemap(fb: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(fb, np.float32).copy()
    h, w = out.shape[:2]
    lib().grrt_tonemap(out.ctypes.data_as(C.POINTER(C.c_float)), w, h)
    return out


def write_fits_cube(path: str, cube: np.ndarray, freqs=None, params: GRRTParams | None = None,
                    bitpix: int = -64) -> None:
    """Write a (height, width, bins) cube as FITS (grrt_ext_write_fits_cube)."""
    dt = np.float64 if bitpix == -64 else np.float32
    c = np.ascontiguousarray(cube, dt)
    h, w, nb = c.shape
    f = None if freqs is None else (C.c_double * nb)(*[float(v) for v in freqs])
    rc = lib().grrt_ext_write_fits_cube(path.encode(), c.ctypes.data, bitpix, w, h, nb, f,
                                        C.byref(params) if params is not None else None)
    if rc != 0:
        raise RuntimeError(lib().grrt_last_error().decode())


def read_texture(path: str) -> np.ndarray:
    """Read a GRRT_BG_TEXTURE image (.hdr or PFM) as (height, width, 3) float32 linear RGB."""
    w, h = C.c_int(0), C.c_int(0)
    if lib().grrt_ext_read_texture(path.encode(), C.byref(w), C.byref(h), None, 0) != 0:
        raise RuntimeError(lib().grrt_last_error().decode())
    out = np.empty((h.value, w.value, 3), np.float32)
    if lib().grrt_ext_read_texture(path.encode(), C.byref(w), C.byref(h),
                                   out.ctypes.data_as(C.POINTER(C.c_float)), out.size) != 0:
        raise RuntimeError(lib().grrt_last_error().decode())
    return out


def write_hdr(path: str, rgba: np.ndarray, normalised: bool = False) -> None:
    """Write an (height, width, 4) float32 frame as a Radiance RGBE .hdr."""
    fb = np.ascontiguousarray(rgba, np.float32)
    h, w, _ = fb.shape
    rc = lib().grrt_ext_write_hdr(path.encode(), fb.ctypes.data_as(C.POINTER(C.c_float)), w, h,
                                  1 if normalised else 0)
    if rc != 0:
        raise RuntimeError(lib().grrt_last_error().decode())
