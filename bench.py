#!/usr/bin/env python3
"""Throughput of the GPU geodesic tracer on the Kerr a=0.998 1024^2 4-spp frame,
plain FP64 and with the learned surrogate (FP64 fallback and judge).

One step = one full frame: 1024 x 1024 pixels x 4 samples (seeded jitter,
seed 1234), thin Novikov–Thorne disk from the ISCO to 20 M, 5000-star
background, observer r=50 M, theta=80 deg, 90 deg field of view, tolerance
1e-8, at most 10000 steps, FP64 state.  The output frame stays in device
memory (the rate including the host copy is reported separately).

Prints one JSON line.  Runs on one GPU; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 1024
SPP = 4
SEED = 1234
FP64_PEAK_TFLOPS = 78.6           # MI355X FP64 vector peak
# Algorithmic FP64 work of one Dormand–Prince attempt in the reference
# formulation: 7 Kerr right-hand sides of ~170 flops plus ~450 flops of stage
# combinations, error norm and control.
ALGO_FLOPS_PER_ATTEMPT = 1640.0
# One classic RK4 ray-march step: 4 right-hand sides + ~120 flops, and the
# transfer of each frequency bin (opacity lookup, source, extinction) ~60.
ALGO_FLOPS_PER_MARCH_STEP = 800.0
ALGO_FLOPS_PER_MARCH_BIN = 60.0
BF16_PEAK_TFLOPS = 2500.0         # MI355X dense BF16 MFMA peak
# Algorithmic work of one surrogate evaluation: 2 * (5*256 + 3*256*256 + 256*9).
SURROGATE_FLOPS_PER_RAY = 2.0 * (5 * 256 + 3 * 256 * 256 + 256 * 9)
# Surrogate gates: class entropy (nats), redshift consistency (ln g), judged share.
SG_ENTROPY, SG_LNG_TOL, SG_JUDGE = 0.1, 0.005, 0.02
# Critical-curve gate (the library's default): rays with |critical score| below it fall back.
SG_CRITICAL = 0.03
# HBM bytes moved by one plain-frame trace launch: rocprofv3 --pmc FETCH_SIZE
# (10 174 KiB, x2 by a copy-kernel calibration) + WRITE_SIZE (118 150 KiB,
# x1), i.e. the 25-byte per-sample outputs, table refills and the few
# registers spilled outside the attempt loop.
TRACE_HBM_BYTES = (10174 * 2 + 118150) * 1024


def kerr_frame_params(grrt):
    return grrt.make_params(width=W, height=H, metric=grrt.METRIC_KERR, spin=0.998,
                            observer_r=50.0, observer_theta=math.radians(80.0),
                            observer_phi=0.0, fov=math.radians(90.0), disk=True,
                            disk_outer=20.0, stars=True, tolerance=1e-8, max_steps=10000,
                            spp=SPP)


def host_cores() -> tuple[int, dict]:
    """Threads for the CPU baseline: the cores this process may run on
    (its affinity mask, further limited by OMP_NUM_THREADS when set, which
    is how the host's share of cores is handed to a job), and a description
    of the host."""
    affinity = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(env)) if env.isdigit() and int(env) > 0 else affinity
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, threads), {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cores": affinity,
                             "omp_num_threads_env": env or None}


def timed_median(fn, runs: int = 3) -> float:
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baseline(target_s: float = 5.0) -> dict:
    """The CPU oracle built with the reference's release flags, OpenMP over
    the host's cores, on a random sample of the same frame's rays; the
    median of three timed runs."""
    import numpy as np
    import oracle as O

    threads, host = host_cores()
    sc = O.scene(width=W, height=H, spin_a=0.998, r_obs=50.0, theta_obs=math.radians(80.0),
                 fov=math.radians(90.0), disk=True, r_out=20.0, stars=True)
    rng = np.random.default_rng(2024)
    n_total = W * H * SPP
    probe = np.sort(rng.choice(n_total, 20000, replace=False))
    t0 = time.perf_counter()
    O.trace_samples(sc, O.JITTER_SEEDED, SPP, SEED, probe, threads=threads, variant="fast")
    rate = len(probe) / (time.perf_counter() - t0)
    n = int(min(n_total, max(20000, rate * target_s)))
    ids = np.sort(rng.choice(n_total, n, replace=False))
    dt = timed_median(lambda: O.trace_samples(sc, O.JITTER_SEEDED, SPP, SEED, ids, threads=threads, variant="fast"))
    return {"value": round(n / dt / 1e6, 5), "unit": "Mrays/s", "cores": threads,
            "kind": "port", "host": host,
            "sample": f"{n} rays drawn uniformly from the same 1024^2x4 frame, median of 3 runs of {dt:.1f} s, "
                      f"oracle built -O3 -ffast-math -mavx2 -mfma -flto, OpenMP guided"}


def pixel_flux_error(a, b, sel=None):
    """Median relative error of per-pixel summed RGB (a against the plain
    frame b) over pixels lit in either frame (and in sel, if given), and the
    share of those pixels within 1%."""
    import numpy as np
    la, lb = a[..., :3].sum(-1).astype(np.float64), b[..., :3].sum(-1).astype(np.float64)
    lit = (la > 0) | (lb > 0)
    if sel is not None:
        lit &= sel.reshape(lit.shape)
    e = np.abs(la[lit] - lb[lit]) / np.where(lb[lit] > 0, lb[lit], np.inf)
    e = np.where(lb[lit] > 0, e, 1.0)
    return float(np.median(e)), float(np.mean(e <= 0.01))


def disk_pixels(*classes):
    """Pixels with a disk-class sample in any of the given per-sample class maps."""
    import numpy as np
    m = None
    for c in classes:
        d = (c.reshape(-1, SPP) == 2).any(1)
        m = d if m is None else (m | d)
    return m


def surrogate_section(grrt, ctx, plain_value, steps, warmup, sg_steps, big_frames):
    """Train the surrogate on rays of the FP64 integrator, then time the
    surrogate + fallback + judge frame on the same workload and compare it
    pixel by pixel with the plain FP64 frame."""
    import numpy as np
    t0 = time.perf_counter()
    ctx.surrogate_init(1)
    ctx.set_critical_gate(SG_CRITICAL)
    ctx.surrogate_generate(1 << 24, seed=7)
    ctx.surrogate_generate(1 << 16, seed=8, validation=True)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tr = ctx.surrogate_train(steps=sg_steps, batch=16384, lr_max=2e-3, lr_min=2e-5, warmup=200,
                             w_r=100.0, w_g=100.0, w_dir=1.0, seed=3)
    train_s = time.perf_counter() - t0
    ev = ctx.surrogate_evaluate()
    plain = ctx.render()
    pcls = ctx.classes()
    fb, _ = ctx.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL)
    scls = ctx.classes()
    med, within = pixel_flux_error(fb, plain, disk_pixels(pcls, scls))
    med_all, _ = pixel_flux_error(fb, plain)
    for _ in range(warmup):
        ctx.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL, device=True)
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = ctx.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL, device=True)
        stats.append(st)
    wall = time.perf_counter() - t0
    rays = W * H * SPP
    ms = wall / steps * 1e3
    value = rays / (ms / 1e3) / 1e6
    infer_ms = sum(s["infer_ms"] for s in stats) / steps
    trace_ms = sum(s["trace_ms"] for s in stats) / steps
    fb_attempts = sum(s["trace_attempts"] for s in stats) / steps
    fb_achieved = ALGO_FLOPS_PER_ATTEMPT * fb_attempts / (trace_ms / 1e3) / 1e12 if trace_ms > 0 else 0.0
    gpu_ms = sum(s["gpu_ms"] for s in stats) / steps
    achieved = SURROGATE_FLOPS_PER_RAY * rays / (infer_ms / 1e3) / 1e12
    last = stats[-1]
    out = {
        "workload": "the same Kerr a=0.998 1024^2 4-spp frame",
        "value": round(value, 3), "unit": "Mrays/s", "ms_per_step": round(ms, 3),
        "speedup_vs_plain": round(value / plain_value, 3),
        "gpu_ms_per_frame": round(gpu_ms, 3), "inference_span_ms": round(infer_ms, 3),
        "trace_after_inference_ms": round(trace_ms, 3),
        "fallback_frac": round(last["fallback_rays"] / last["rays"], 5),
        "judged_frac": round(last["judged_rays"] / last["rays"], 5),
        "class_agreement_vs_plain": round(float(np.mean(scls == pcls)), 6),
        "median_disk_pixel_flux_rel_err_vs_plain": round(med, 5),
        "disk_pixels_within_1pct": round(within, 4),
        "median_lit_pixel_flux_rel_err_vs_plain": round(med_all, 5),
        "judge_class_agreement": round(last["judge_class_agreement"], 6),
        "judge_median_rel_err": round(last["judge_median_rel_err"], 5),
        "gates": {"entropy_max_nats": SG_ENTROPY, "lng_tol": SG_LNG_TOL, "judge_fraction": SG_JUDGE,
                  "critical_score_max": SG_CRITICAL},
        "training": {"rays": 1 << 24, "steps": sg_steps, "batch": 16384, "generate_s": round(gen_s, 2),
                     "train_s": round(train_s, 2), "val_accuracy": round(ev["accuracy"], 5),
                     "val_rmse_ln_r": round(ev["rmse_lnr"], 5), "val_rmse_ln_g": round(ev["rmse_lng"], 5)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": None},
        # the FP64 fallback/judge trace after the network pass (listed rays)
        "fallback_roofline": {"bound": "valu", "achieved": round(fb_achieved, 3), "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": round(fb_achieved / FP64_PEAK_TFLOPS, 4),
                              "attempts": int(fb_attempts), "traffic": None},
    }
    if big_frames > 0:
        # The surrogate's own inference frame: Kerr a=0.998, theta=80 deg,
        # 2048^2 x 4 spp (no star background named there).
        pbig = kerr_frame_params(grrt)
        pbig.width = pbig.height = 2048
        pbig.background_type = 0
        with grrt.Context(pbig) as cbig:
            cbig.set_sampling(grrt.SAMPLING_SEEDED, SPP, SEED)
            cbig.surrogate_init(1)
            cbig.set_critical_gate(SG_CRITICAL)
            cbig.surrogate_params(ctx.surrogate_params())
            plain_big = cbig.render()
            plain_big_cls = cbig.classes()
            cbig.render_device()
            t0 = time.perf_counter()
            for _ in range(big_frames):
                cbig.render_device()
            plain_ms = (time.perf_counter() - t0) / big_frames * 1e3
            f_big, _ = cbig.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL)
            sg_big_cls = cbig.classes()
            med_big, within_big = pixel_flux_error(f_big, plain_big, disk_pixels(plain_big_cls, sg_big_cls))
            cbig.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL, device=True)
            t0 = time.perf_counter()
            big_stats = []
            for _ in range(big_frames):
                _, st_big = cbig.render_surrogate(SG_ENTROPY, SG_JUDGE, SG_LNG_TOL, device=True)
                big_stats.append(st_big)
            sg_ms = (time.perf_counter() - t0) / big_frames * 1e3
        rays_big = 2048 * 2048 * SPP
        out["frame_2048sq_4spp"] = {
            "plain_mrays_s": round(rays_big / plain_ms / 1e3, 3), "surrogate_mrays_s": round(rays_big / sg_ms / 1e3, 3),
            "speedup": round(plain_ms / sg_ms, 3), "plain_ms": round(plain_ms, 3), "surrogate_ms": round(sg_ms, 3),
            "class_agreement_vs_plain": round(float(np.mean(sg_big_cls == plain_big_cls)), 6),
            "median_disk_pixel_flux_rel_err_vs_plain": round(med_big, 5),
            "fallback_frac": round(st_big["fallback_rays"] / st_big["rays"], 5),
            "judge_median_rel_err": round(st_big["judge_median_rel_err"], 5),
            "gpu_ms_per_frame": round(sum(x["gpu_ms"] for x in big_stats) / big_frames, 3),
            "inference_span_ms": round(sum(x["infer_ms"] for x in big_stats) / big_frames, 3),
            "trace_after_inference_ms": round(sum(x["trace_ms"] for x in big_stats) / big_frames, 3)}
    return out


def uneven_views(seed: int = 7, n: int = 64):
    """The batched-views workload: n thin-disk Kerr views drawn from one
    PRNG stream in a fixed order per view (a, theta, phi, log10 r, fov,
    resolution): a ~ U[0, 0.998], theta ~ U[5, 89] deg, phi ~ U[0, 360) deg,
    r log-uniform in [20, 100], fov ~ U[30, 90] deg, square side from
    {128, 256, 512, 1024}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(0.0, 0.998)
        th = math.radians(rng.uniform(5.0, 89.0))
        ph = math.radians(rng.uniform(0.0, 360.0))
        r = 10.0 ** rng.uniform(math.log10(20.0), math.log10(100.0))
        fov = math.radians(rng.uniform(30.0, 90.0))
        side = int(rng.choice([128, 256, 512, 1024]))
        out.append((a, r, th, ph, fov, side, side))
    return out


def batched_views_section(grrt, reps: int) -> dict:
    """64 uneven views in one persistent launch (recorded graph, autotuned
    launch shape) against the same views rendered one after another."""
    views = uneven_views()
    rays = sum(v[5] * v[6] for v in views)
    base = grrt.make_params(width=64, height=64, metric=grrt.METRIC_KERR, spin=0.5, disk=True, disk_outer=20.0,
                            stars=False, tolerance=1e-8, max_steps=10000)
    with grrt.Context(base) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        best, tune_ms = ctx.autotune(views)
        ctx.render_batch(views, device=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.render_batch(views, device=True)
        batch_ms = (time.perf_counter() - t0) / reps * 1e3
        kernel_ms = ctx.stats().trace_ms
        batch_attempts = ctx.stats().attempts
    # The same views one at a time: a context per view (created outside the
    # timed loop), each frame left in device memory like the batch's.
    ctxs = []
    for a, r, th, ph, fov, w, h in views:
        p = grrt.make_params(width=w, height=h, metric=grrt.METRIC_KERR, spin=a, observer_r=r, observer_theta=th,
                             observer_phi=ph, fov=fov, disk=True, disk_outer=20.0, stars=False, tolerance=1e-8,
                             max_steps=10000)
        c = grrt.Context(p)
        c.set_sampling(grrt.SAMPLING_CENTER)
        ctxs.append(c)
    for c in ctxs[:2]:
        c.render_device()
    t0 = time.perf_counter()
    for c in ctxs:
        c.render_device()
    seq_ms = (time.perf_counter() - t0) * 1e3
    for c in ctxs:
        c.close()
    return {"workload": "64 thin-disk Kerr views from seed 7 (a, theta, phi, r, fov, side from "
                        "{128..1024}), pixel centres, tol 1e-8",
            "rays": rays, "batch_mrays_s": round(rays / batch_ms / 1e3, 3), "batch_ms": round(batch_ms, 3),
            "batch_kernel_ms": round(kernel_ms, 3), "autotuned": best, "autotune_best_ms": round(tune_ms, 3),
            "roofline": {"bound": "valu",
                         "achieved": round(ALGO_FLOPS_PER_ATTEMPT * batch_attempts / (kernel_ms / 1e3) / 1e12, 3),
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(ALGO_FLOPS_PER_ATTEMPT * batch_attempts / (kernel_ms / 1e3) / 1e12
                                       / FP64_PEAK_TFLOPS, 4), "traffic": None},
            "sequential_mrays_s": round(rays / seq_ms / 1e3, 3), "sequential_ms": round(seq_ms, 3),
            "batch_speedup": round(seq_ms / batch_ms, 3)}


CUBE_FREQS = [10 ** (14 + 4 * k / 31) for k in range(32)]
# The cube's longest ray (pixel x, y): 4057 Dormand-Prince attempts and 3777
# ray-march steps over 3771 entries into the disk volume (CPU oracle).
LONGEST_CUBE_RAY = (274, 253)


def spectral_cube_params(grrt):
    return grrt.make_params(width=512, height=512, metric=grrt.METRIC_KERR, spin=0.9,
                            observer_theta=math.radians(75.0), fov=math.radians(90.0), volumetric=True,
                            disk_outer=20.0, turbulence=0.4, seed=42, octaves=2, alpha=0.1,
                            freqs=CUBE_FREQS)


def spectral_cube_section(grrt, reps: int, cpu: bool) -> dict:
    """The volumetric spectral cube: Kerr a=0.9, 512^2 x 1 spp (pixel-hash
    jitter), theta=75 deg, alpha=0.1, turbulence 0.4, 2 octaves, seed 42,
    32 frequencies log-spaced over 1e14-1e18 Hz, float32 cube; timed in the
    mixed form (FP32 transfer, FP64 geodesic) and all-FP64, with the CPU
    oracle (all FP64) on a random sample of the same rays."""
    p = spectral_cube_params(grrt)
    t0 = time.perf_counter()
    ctx = grrt.Context(p)
    setup_s = time.perf_counter() - t0
    ctx.set_frequency_bins(CUBE_FREQS)
    rays = 512 * 512
    out = {"workload": "Kerr a=0.9 512^2 x 1 spp, theta=75deg, volumetric alpha=0.1 turb 0.4 2 oct seed 42, "
                       "32 bins 1e14-1e18 Hz, float32 cube",
           "context_setup_s": round(setup_s, 3)}
    # Set-up of the disk tables: opacity tables and vertical columns on host
    # threads against the GPU (grrt_create's path), the radial structure on
    # the host in both; best of three builds each.
    builds = {}
    for dev in (False, True):
        runs = [grrt.volumetric_tables(p, device=dev) for _ in range(3)]
        builds["device" if dev else "host"] = (min(r["build_ms_opacity"] for r in runs),
                                               min(r["build_ms_total"] for r in runs),
                                               min(r["build_ms_columns"] for r in runs),
                                               min(r["build_ms_columns_kernel"] for r in runs))
    out["table_build_ms"] = {"opacity_host": round(builds["host"][0], 2),
                             "opacity_device": round(builds["device"][0], 2),
                             "columns_host": round(builds["host"][2], 2),
                             "columns_device": round(builds["device"][2], 2),
                             "columns_device_kernel": round(builds["device"][3], 2),
                             "total_host": round(builds["host"][1], 2),
                             "total_device": round(builds["device"][1], 2),
                             # the library's own table build uses min(16, hardware threads)
                             "host_threads": max(1, min(16, os.cpu_count() or 1))}
    for mode, name in ((1, "mixed"), (0, "fp64"), (2, "mixed_far_field_fp32")):
        ctx.set_precision(mode)
        ctx.render_spectral_f32()
        wall, kern = [], []
        for _ in range(reps):
            t1 = time.perf_counter()
            ctx.render_spectral_f32()
            wall.append(time.perf_counter() - t1)
            kern.append(ctx.stats().trace_ms)
        ms = sum(wall) / reps * 1e3
        st = ctx.stats()
        k_ms = sum(kern) / reps
        flops = (ALGO_FLOPS_PER_ATTEMPT * st.attempts +
                 (ALGO_FLOPS_PER_MARCH_STEP + ALGO_FLOPS_PER_MARCH_BIN * len(CUBE_FREQS)) * st.march_steps)
        achieved = flops / (k_ms / 1e3) / 1e12
        out[name] = {"ms_per_cube": round(ms, 2), "trace_kernel_ms": round(k_ms, 2),
                     "mrays_s": round(rays / (ms / 1e3) / 1e6, 3),
                     "attempts_per_cube": int(st.attempts), "march_steps_per_cube": int(st.march_steps),
                     "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None}}
    # The cube is bound by its longest ray: the same ray traced alone.
    ctx.set_precision(0)
    tile = []
    for _ in range(3):
        ctx.render_spectral_tile(LONGEST_CUBE_RAY[0], LONGEST_CUBE_RAY[1], 1, 1)
        tile.append(ctx.stats().trace_ms)
    out["longest_ray_alone_ms"] = round(sorted(tile)[1], 2)
    out["longest_ray"] = {"pixel": list(LONGEST_CUBE_RAY), "attempts": int(ctx.stats().attempts),
                          "march_steps": int(ctx.stats().march_steps)}
    ctx.close()
    if cpu:
        import numpy as np
        import oracle as O
        T = grrt.volumetric_tables(p)
        V = O.vol_struct(T)
        sc = O.scene(width=512, height=512, spin_a=0.9, theta_obs=math.radians(75.0), fov=math.radians(90.0),
                     disk=False)
        threads, host = host_cores()
        rng = np.random.default_rng(2025)
        probe = np.sort(rng.choice(rays, 1000, replace=False))
        t1 = time.perf_counter()
        O.vol_trace_samples(sc, V, O.JITTER_HASH, 1, 0, probe, freqs=CUBE_FREQS, threads=threads, variant="fast")
        rate = len(probe) / (time.perf_counter() - t1)
        n = int(min(rays, max(1000, rate * 4.0)))
        ids = np.sort(rng.choice(rays, n, replace=False))
        dt = timed_median(lambda: O.vol_trace_samples(sc, V, O.JITTER_HASH, 1, 0, ids, freqs=CUBE_FREQS,
                                                      threads=threads, variant="fast"))
        cube_s = rays / (n / dt)
        out["cpu_baseline"] = {"value": round(n / dt / 1e6, 5), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "cube_s": round(cube_s, 2), "host": host,
                               "sample": f"{n} rays drawn uniformly from the cube's 512^2, median of 3 runs of "
                                         f"{dt:.1f} s, oracle built -O3 -ffast-math, all FP64, OpenMP guided"}
        out["speedup_mixed_vs_cpu"] = round(cube_s * 1e3 / out["mixed"]["ms_per_cube"], 1)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", default="", help="block,blocks_per_cu,refill_min,chunk[,tile]")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sg-steps", type=int, default=60000, help="surrogate training steps")
    ap.add_argument("--big-frames", type=int, default=4, help="2048^2 frames timed per form (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=2, help="replays of the 64-view batch (0: skip)")
    ap.add_argument("--cube-reps", type=int, default=3, help="volumetric spectral cubes timed per form (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              f"(rays are independent: each GPU would trace its own interleaved tiles)",
              file=sys.stderr)
        return 2

    import grrt_amd as grrt

    params = kerr_frame_params(grrt)
    ctx = grrt.Context(params)
    ctx.set_sampling(grrt.SAMPLING_SEEDED, SPP, SEED)
    if args.tune:
        t = [int(x) for x in args.tune.split(",")]
        ctx.set_tuning(t[0], t[1], t[2], t[3], t[4] if len(t) > 4 else -1)
    for _ in range(args.warmup):
        ctx.render_device()
    trace_ms, attempts, steps, derivs = [], 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.render_device()          # returns after the frame is complete
        s = ctx.stats()
        trace_ms.append(s.trace_ms)
        attempts += s.attempts
        steps += s.steps
        derivs += s.deriv_evals
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    rays = W * H * SPP
    value = rays / (ms_per_step / 1e3) / 1e6
    avg_trace_s = sum(trace_ms) / len(trace_ms) / 1e3
    att_per_launch = attempts / args.steps
    achieved = ALGO_FLOPS_PER_ATTEMPT * att_per_launch / avg_trace_s / 1e12
    # Rate including the copy to a host framebuffer (grrt_render into the
    # caller's buffer, reused as a renderer would): median of five after a
    # warm-up call.
    import numpy as np
    host_fb = np.zeros((H, W, 4), np.float32)
    ctx.render(out=host_fb)
    host_runs, copy_runs, host_trace = [], [], []
    for _ in range(5):
        t1 = time.perf_counter()
        ctx.render(out=host_fb)
        host_runs.append((time.perf_counter() - t1) * 1e3)
        copy_runs.append(ctx.stats().copy_ms)
        host_trace.append(ctx.stats().trace_ms)
    host_ms = sorted(host_runs)[2]
    # wall time of each host render over its own trace kernel's time
    host_over_trace = sorted(h / t for h, t in zip(host_runs, host_trace))[2]
    # device tone mapping of the frame (grrt_ext_tonemap_device): 16 B read
    # in the statistics pass, 16 B read and 12 B written in the mapping pass
    tm_ms = []
    for _ in range(3):
        ctx.render_device()
        tm_ms.append(ctx.tonemap_device())
    tm_ms = sorted(tm_ms)[1]
    tm_gbs = W * H * 44 / (tm_ms / 1e3) / 1e9

    out = {
        "metric": "Mrays/s Kerr a=0.998 1024\u00b2 4spp: plain FP64 vs surrogate @\u22641% flux err; %VALU pk",
        # value: the plain FP64 form; the surrogate form is under "surrogate"
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: camera rays of the Kerr a=0.998 thin-disk scene",
        "config": {"workload": "Kerr a=0.998 M=1, 1024x1024, 4 spp seeded jitter (seed 1234), "
                               "r_obs=50 theta=80deg fov=90deg, thin disk r_isco..20M, stars, "
                               "tol 1e-8, max 10000 steps",
                   "rays_per_step": rays, "parallelism": "single GPU"},
        "geodesic_steps_per_s": round(steps / args.steps / (ms_per_step / 1e3), 1),
        "attempts_per_s": round(attempts / args.steps / (ms_per_step / 1e3), 1),
        "derivative_evals_per_s": round(derivs / args.steps / (ms_per_step / 1e3), 1),
        "attempts_per_launch": int(att_per_launch),
        "trace_kernel_ms": round(avg_trace_s * 1e3, 3),
        "host_copy_inclusive_ms": round(host_ms, 3),
        "host_copy_inclusive_mrays_s": round(rays / host_ms / 1e3, 3),
        "host_copy_ms": round(sorted(copy_runs)[2], 3),
        "host_copy_inclusive_over_trace": round(host_over_trace, 4),
        "tonemap_device": {"gpu_ms": round(tm_ms, 4), "algorithmic_gb_s": round(tm_gbs, 1),
                           "hbm_frac": round(tm_gbs / 8000.0, 4), "bytes_per_pixel": 44},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": TRACE_HBM_BYTES},
    }
    if not args.no_surrogate:
        out["surrogate"] = surrogate_section(grrt, ctx, value, args.steps, args.warmup, args.sg_steps,
                                             args.big_frames)
    ctx.close()
    if args.batch_reps > 0:
        out["batched_views"] = batched_views_section(grrt, args.batch_reps)
    if args.cube_reps > 0:
        out["volumetric_spectral"] = spectral_cube_section(grrt, args.cube_reps, not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
