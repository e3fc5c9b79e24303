"""The learned surrogate on the GPU: the MFMA forward pass and the backward
pass against the CPU oracle of the same network (oracle/mlp_oracle.cpp, with
the GPU's BF16 rounding points emulated), the training targets against the
oracle integrator, training progress, and the surrogate frame (fallback rays
are the integrator's own results; judged rays report the surrogate error).

Tolerances: the GPU accumulates in FP32 and evaluates SiLU in FP32, the
oracle in FP64, so an activation next to a BF16 rounding boundary can round
the other way; 99% of outputs agree to 1e-4 relative, gradients to 1e-2
relative per layer; a few outputs in a thousand differ by up to 1e-2 where a
hidden activation rounded the other way (measured agreement is printed).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(256, 5), (256, 256), (256, 256), (256, 256), (9, 256)]


def _ctx(grrt, **kw):
    p = grrt.make_params(**{**dict(width=64, height=64, metric=1, spin=0.998, disk=True, disk_outer=20.0), **kw})
    return grrt.Context(p)


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    t = np.zeros((n, 8), np.float32)
    t[:, 0] = rng.integers(0, 3, n)
    t[: n // 16, 0] = -1
    t[:, 1] = rng.uniform(0.2, 3.0, n)
    t[:, 2] = rng.uniform(-1.0, 0.4, n)
    th, ph = rng.uniform(0, np.pi, n), rng.uniform(-np.pi, np.pi, n)
    t[:, 3:7] = np.stack([np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)], 1)
    t[:, 7] = rng.integers(0, 2, n)
    return x, t


def test_forward_matches_oracle(grrt, oracle):
    print("case: surrogate forward 1000 rows vs oracle (BF16 emulated)", flush=True)
    with _ctx(grrt) as ctx:
        ctx.surrogate_init(123)
        p = ctx.surrogate_params()
        x, _ = _rows(1000, 1)
        y = ctx.surrogate_forward(x)
    ref = oracle.mlp_forward(p, x, emulate=True)
    err = np.abs(y - ref) / (np.abs(ref) + 0.1)
    print(f"forward: max {err.max():.2e} median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e}", flush=True)
    assert err.max() <= 2e-2                      # an activation rounded the other way
    assert np.percentile(err, 99) <= 1e-4
    assert np.median(err) <= 1e-5                 # hi + lo carries ~16 bits of the last layer


def test_gradients_match_oracle(grrt, oracle):
    print("case: surrogate gradients 512 rows vs oracle (BF16 emulated)", flush=True)
    w = (3.0, 2.0, 0.5)
    x, t = _rows(512, 2)
    with _ctx(grrt) as ctx:
        ctx.surrogate_init(7)
        p = ctx.surrogate_params()
        loss, g = ctx.surrogate_gradients(x, t, w)
    oloss, og = oracle.mlp_gradients(p, x, t, w, emulate=True)
    assert loss == pytest.approx(oloss, rel=1e-4)
    o = 0
    for li, (n, k) in enumerate(SIZES):
        for name, cnt in (("W", n * k), ("b", n)):
            a, b = g[o:o + cnt].astype(np.float64), og[o:o + cnt]
            o += cnt
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            print(f"layer {li} {name}: rel {rel:.2e} cos {cos:.8f}", flush=True)
            assert rel <= 1e-2 and cos >= 0.9999, (li, name, rel, cos)


def test_training_targets_match_oracle_integrator(grrt, oracle):
    """Rows of the generated training set against the oracle tracing the same
    ray (camera parameters recovered from the stored features)."""
    print("case: surrogate training rows vs oracle trace", flush=True)
    with _ctx(grrt) as ctx:
        ctx.surrogate_init(1)
        ctx.surrogate_generate(1 << 14, seed=5)
        x, t = ctx.surrogate_dataset(1 << 14)
    rng = np.random.default_rng(3)
    rows = rng.choice(1 << 14, 300, replace=False)
    deg = math.pi / 180
    clo, chi = math.cos(89 * deg), math.cos(5 * deg)
    agree, dr, dg = 0, [], []
    for i in rows:
        f = x[i].astype(np.float64)
        alpha, beta = f[0] * math.pi / 4, f[1] * math.pi / 4
        a = (f[2] + 1) / 2 * 0.998
        th = math.acos((f[3] + 1) / 2 * (chi - clo) + clo)
        r = math.exp((f[4] + 1) / 2 * math.log(5) + math.log(20))
        sc = oracle.scene(width=32, height=32, spin_a=a, r_obs=r, theta_obs=th, fov=math.pi / 2, disk=True,
                          r_out=20.0)
        px, py = alpha * 32 / (math.pi / 2) + 16, beta * 32 / (math.pi / 2) + 16
        ray = oracle.trace(sc, oracle.camera_ray(sc, px, py))
        cls = -1 if ray.cls == 3 else ray.cls
        if cls == int(t[i, 0]):
            agree += 1
            if cls == 2:
                dr.append(abs(math.log(ray.r_hit) - t[i, 1]))
                dg.append(abs(math.log(ray.g_hit) - t[i, 2]))
    print(f"class agreement {agree}/300, median |d ln r| {np.median(dr):.1e}, |d ln g| {np.median(dg):.1e}",
          flush=True)
    assert agree >= 294
    assert np.median(dr) <= 1e-4 and np.median(dg) <= 1e-4


def test_training_reduces_validation_loss(grrt):
    print("case: surrogate training 2^18 rays, 400 steps", flush=True)
    with _ctx(grrt) as ctx:
        ctx.surrogate_init(11)
        ctx.surrogate_generate(1 << 18, seed=21)
        ctx.surrogate_generate(1 << 15, seed=22, validation=True)
        before = ctx.surrogate_evaluate()
        st = ctx.surrogate_train(steps=400, batch=2048, lr_max=2e-3, lr_min=1e-4, warmup=40)
        after = ctx.surrogate_evaluate()
    print("before", before, "\ntrain", st, "\nafter", after, flush=True)
    assert after["loss"] < 0.3 * before["loss"]
    assert after["accuracy"] > 0.9


def _err_stats(y, ref):
    err = np.abs(y - ref) / (np.abs(ref) + 0.1)
    return err, float(np.median(err)), float(np.percentile(err, 99)), float(err.max())


def test_inference_pass_matches_oracle_on_inputs(grrt, oracle):
    """mlp_infer_kernel itself (pre-permuted BF16 weights streamed through
    the LDS windows, hi + lo input and output layers) on 2^14 fixed inputs,
    against the oracle with the GPU's rounding points: median 1e-5, p99 1e-4."""
    print("case: surrogate inference pass, 16384 rows on fixed weights vs oracle", flush=True)
    x = np.random.default_rng(9).uniform(-1.1, 1.1, (16384, 5)).astype(np.float32)
    with _ctx(grrt) as ctx:
        ctx.surrogate_init(123)
        p = ctx.surrogate_params()
        y = ctx.surrogate_infer_raw(x)
        y2 = ctx.surrogate_forward(x)
    ref = oracle.mlp_forward(p, x, emulate=True)
    _, med, p99, mx = _err_stats(y, ref)
    print(f"inference pass: median {med:.2e} p99 {p99:.2e} max {mx:.2e}", flush=True)
    assert med <= 1e-5 and p99 <= 1e-4 and mx <= 2e-2
    # the training forward kernel and the inference pass run the same network
    _, med2, p992, _ = _err_stats(y, y2.astype(np.float64))
    assert med2 <= 1e-5 and p992 <= 1e-4


def test_inference_pass_inputs_from_the_view(grrt, oracle):
    """The pass computing its own inputs from each sample's pixel position
    (camera angles of src/camera.cpp:69-96, scaled to the training box),
    against the oracle on the same inputs computed here."""
    print("case: surrogate inference pass, inputs from a 128x96 centre-sampled view", flush=True)
    W, H = 128, 96
    th, fov, a, r = math.radians(80.0), math.radians(90.0), 0.998, 50.0
    with _ctx(grrt, width=W, height=H, observer_theta=th, fov=fov) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.surrogate_init(5)
        p = ctx.surrogate_params()
        y = ctx.surrogate_infer_raw(None)
    jj, ii = np.divmod(np.arange(W * H), W)
    alpha = (ii + 0.5 - W / 2.0) * fov / W
    beta = (jj + 0.5 - H / 2.0) * fov / W
    clo, chi = math.cos(math.radians(89)), math.cos(math.radians(5))
    x = np.stack([alpha / (math.pi / 4), beta / (math.pi / 4), np.full(W * H, 2 * a / 0.998 - 1),
                  np.full(W * H, 2 * (math.cos(th) - clo) / (chi - clo) - 1),
                  np.full(W * H, 2 * (math.log(r) - math.log(20)) / math.log(5) - 1)], 1).astype(np.float32)
    ref = oracle.mlp_forward(p, x, emulate=True)
    _, med, p99, mx = _err_stats(y, ref)
    print(f"view inputs: median {med:.2e} p99 {p99:.2e} max {mx:.2e}", flush=True)
    assert med <= 1e-5 and p99 <= 1e-4 and mx <= 2e-2


def test_adam_kernel_matches_oracle(grrt, oracle):
    """The training loop's Adam kernel against oracle.mlp_adam (itself pinned
    to torch.optim.Adam) over three consecutive steps (bias corrections of
    t = 1, 2, 3)."""
    print("case: Adam kernel, 3 steps on 201225 parameters vs oracle", flush=True)
    rng = np.random.default_rng(17)
    n = 201225
    w0 = rng.normal(0, 0.05, n).astype(np.float32)
    m0 = rng.normal(0, 1e-3, n).astype(np.float32)
    v0 = np.abs(rng.normal(0, 1e-5, n)).astype(np.float32)
    wg, mg, vg = w0.copy(), m0.copy(), v0.copy()
    wo, mo, vo = w0.copy(), m0.copy(), v0.copy()
    with _ctx(grrt) as ctx:
        for t in (1, 2, 3):
            g = rng.normal(0, 1e-2, n).astype(np.float32)
            g[::97] = 0.0
            ctx.surrogate_adam_step(wg, g, mg, vg, 1e-3, t)
            oracle.mlp_adam(wo, g, mo, vo, 1e-3, t)
    # m and v: FP32 moment updates; the GPU may fuse b * m + (1 - b) g into
    # one FMA where the oracle rounds twice, so m (a difference of terms of
    # either sign) is compared against the size of its terms.
    for name, a, b in (("m", mg, mo), ("v", vg, vo)):
        err = np.abs(a.astype(np.float64) - b)
        scale = np.abs(b).max()
        print(f"{name}: max abs err {err.max():.2e} (scale {scale:.2e})", flush=True)
        assert np.all(err <= 1e-6 * np.abs(b) + 1e-6 * scale)
    du_g, du_o = wg.astype(np.float64) - w0, wo.astype(np.float64) - w0
    rel = np.abs(du_g - du_o) / np.maximum(np.abs(du_o), 1e-7)
    print(f"w update: median rel {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    assert np.median(rel) <= 1e-5
    # every element: the FP32 update to ~1e-5 of itself, plus the rounding of w
    ulp = np.spacing(np.abs(wo)).astype(np.float64)
    # (three steps: three roundings of w, plus the moments' contraction)
    viol = np.abs(du_g - du_o) / (1e-5 * np.abs(du_o) + ulp)
    for i in np.argsort(-viol)[:3]:
        print(f"worst {i}: ratio {viol[i]:.2f} du_gpu {du_g[i]:.6e} du_oracle {du_o[i]:.6e} w0 {w0[i]:.3e} "
              f"m0 {m0[i]:.3e} v0 {v0[i]:.3e} m {mo[i]:.3e} v {vo[i]:.3e}", flush=True)
    assert viol.max() <= 4.0


def _judge_from_plain(ids, scls, pcls, srgb, prgb):
    """The judge's statistics recomputed from the plain frame: class
    agreement and the median relative error of the summed RGB over lit rays."""
    agree = float(np.mean(scls[ids] == pcls[ids])) if len(ids) else 1.0
    ls, lr = srgb[ids].sum(1), prgb[ids].sum(1)
    e = np.where(lr > 0, np.abs(ls - lr) / np.where(lr > 0, lr, 1.0), np.where(ls > 0, 1.0, -1.0))
    lit = e[e >= 0]
    return agree, (float(np.sort(lit.astype(np.float32))[len(lit) // 2]) if len(lit) else 0.0)


def test_surrogate_frame_fallback_and_judge(grrt):
    """Short training, then a 128^2 frame: every fallback ray carries the plain
    integrator's class and colour exactly, and the judge's statistics equal
    the ones recomputed from the plain frame."""
    print("case: surrogate frame 128^2 a=0.998 after short training", flush=True)
    with _ctx(grrt, width=128, height=128, observer_theta=math.radians(80.0), fov=math.radians(90.0)) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        plain = ctx.render()
        pcls = ctx.classes()
        prgb = ctx.samples()
        ctx.surrogate_init(3)
        ctx.surrogate_generate(1 << 18, seed=31)
        ctx.surrogate_train(steps=300, batch=2048, lr_max=2e-3, lr_min=1e-4, warmup=30)
        fb, st = ctx.render_surrogate(entropy_max=0.1, judge_fraction=0.05)
        scls = ctx.classes()
        srgb = ctx.samples()
        fbl, jl = ctx.surrogate_lists(st["fallback_rays"], st["judged_rays"])
    print(st, flush=True)
    n = 128 * 128
    assert st["rays"] == n
    assert 0 < st["judged_rays"] < 0.1 * n
    assert len(np.unique(fbl)) == len(fbl) and len(np.unique(jl)) == len(jl)
    assert not np.isin(jl, fbl).any()                 # judged rays are shaded ones
    assert np.array_equal(scls[fbl], pcls[fbl])
    assert np.array_equal(srgb[fbl], prgb[fbl])
    agree = np.mean(scls == pcls)
    print(f"class agreement vs plain {agree:.4f}", flush=True)
    assert agree > 0.9
    ja, jm = _judge_from_plain(jl, scls, pcls, srgb, prgb)
    assert st["judge_class_agreement"] == pytest.approx(ja, abs=1e-12)
    assert st["judge_median_rel_err"] == pytest.approx(jm, rel=1e-5, abs=1e-7)
    assert np.isfinite(fb).all() and np.all(fb[..., 3] == 1.0)
    # With the threshold below 0 nats every ray falls back: the frame is the
    # integrator's, sample for sample.
    with _ctx(grrt, width=128, height=128, observer_theta=math.radians(80.0), fov=math.radians(90.0)) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.surrogate_init(3)
        fb0, st0 = ctx.render_surrogate(entropy_max=-1.0, judge_fraction=0.0)
        assert st0["fallback_rays"] == 128 * 128 and st0["judged_rays"] == 0
        assert np.array_equal(ctx.classes(), pcls)
    assert np.array_equal(fb0, plain)


def test_critical_gate_sends_near_critical_rays_to_the_integrator(grrt, oracle):
    """With the entropy and redshift gates off, the fallback list of a frame
    zoomed on the shadow is exactly the rays whose critical score (the CPU
    restatement, brute-force minimum over 4001 radii) is below the gate,
    apart from rays within 1e-4 of it (the sampling error of the
    restatement, ~1e-6 here); the gate off lists no ray; the listed rays
    carry the integrator's results bit for bit."""
    from parity_util import critical_score
    print("case: critical-curve gate, 128^2 a=0.998 fov 20 deg, gate 0.03", flush=True)
    W, a, gate = 128, 0.998, 0.03
    kw = dict(width=W, height=W, observer_theta=math.radians(80.0), fov=math.radians(20.0), stars=False)
    with _ctx(grrt, **kw) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.render()
        pcls, prgb = ctx.classes(), ctx.samples()
        ctx.surrogate_init(5)
        ctx.set_critical_gate(0.0)
        _, st0 = ctx.render_surrogate(entropy_max=1e30, judge_fraction=0.0)
        ctx.set_critical_gate(gate)
        _, st = ctx.render_surrogate(entropy_max=1e30, judge_fraction=0.0)
        fbl, _ = ctx.surrogate_lists(st["fallback_rays"], 0)
        scls, srgb = ctx.classes(), ctx.samples()
    assert st0["fallback_rays"] == 0
    sc = oracle.scene(width=W, height=W, spin_a=a, theta_obs=math.radians(80.0), fov=math.radians(20.0))
    cam = np.array([oracle.camera_ray(sc, (k % W) + 0.5, (k // W) + 0.5) for k in range(W * W)])
    score = np.abs(critical_score(a, 1.0, cam[:, 4], cam[:, 6], cam[:, 7], cam[:, 2]))
    listed = np.zeros(W * W, bool)
    listed[fbl] = True
    must, may = score < gate - 1e-4, score < gate + 1e-4
    print(f"listed {listed.sum()}, below the gate {np.sum(score < gate)}, firm {must.sum()}", flush=True)
    assert must.sum() > 100
    assert np.all(listed[must]) and not np.any(listed & ~may)
    assert np.array_equal(scls[fbl], pcls[fbl]) and np.array_equal(srgb[fbl], prgb[fbl])


def test_surrogate_texture_background(grrt, tmp_path):
    """A constant texture: every escaped sample, shaded at the network's
    predicted direction or traced by the integrator, carries exactly that
    colour (in float); captured samples stay black."""
    from parity_util import write_pfm
    print("case: surrogate frame 96^2 a=0.998 with a constant texture", flush=True)
    colour = np.float32([0.25, 0.5, 1.0])
    path = str(tmp_path / "const.pfm")
    write_pfm(path, np.broadcast_to(colour, (4, 8, 3)))
    kw = dict(width=96, height=96, observer_theta=math.radians(80.0), fov=math.radians(90.0))
    with _ctx(grrt, texture=path, **kw) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.surrogate_init(3)
        ctx.surrogate_generate(1 << 17, seed=31)
        ctx.surrogate_train(steps=200, batch=2048, lr_max=2e-3, lr_min=1e-4, warmup=20)
        fb, st = ctx.render_surrogate(entropy_max=0.1, judge_fraction=0.0)
        cls, rgb = ctx.classes(), ctx.samples()
    esc = cls == grrt.CLASS_ESCAPED
    print(f"escaped {esc.sum()} of {cls.size}, fallback {st['fallback_rays']}", flush=True)
    assert esc.sum() > 1000
    assert np.array_equal(rgb[esc].astype(np.float32), np.broadcast_to(colour, (esc.sum(), 3)))
    assert np.all(rgb[cls == grrt.CLASS_CAPTURED] == 0.0)
    assert np.isfinite(fb).all()


def _disk_pixel_flux_err(s_fb, p_fb, s_cls, p_cls, nS):
    disk = (s_cls.reshape(-1, nS) == 2).any(1) | (p_cls.reshape(-1, nS) == 2).any(1)
    a = s_fb.reshape(-1, 4)[disk, :3].sum(1).astype(np.float64)
    b = p_fb.reshape(-1, 4)[disk, :3].sum(1).astype(np.float64)
    sel = b > 0
    return float(np.median(np.abs(a[sel] - b[sel]) / b[sel]))


def test_surrogate_2048_4spp_after_full_training(grrt):
    """The learned-surrogate workload at its full size: trained on 2^24 rays
    of the FP64 integrator (60 000 Adam steps of 16 384), then the Kerr
    a=0.998 theta=80 deg frame at 2048^2 x 4 spp with the 0.1-nat entropy
    gate, the redshift-consistency gate and a 2% judge.  Fallback samples
    equal the plain frame's bit for bit, classes agree on >= 99.5% of
    samples, the judge's statistics are recomputed from the plain frame, the
    median disk-pixel flux error is <= 1%, and the trained network makes the
    frame faster than the plain FP64 frame (what it gains over the method
    without it)."""
    print("case: surrogate 2048^2 x 4 spp after 2^24-ray training", flush=True)
    import time
    W = 2048
    with _ctx(grrt, width=W, height=W, observer_theta=math.radians(80.0), fov=math.radians(90.0),
              stars=False) as ctx:
        ctx.set_sampling(grrt.SAMPLING_SEEDED, 4, 1234)
        ctx.surrogate_init(1)
        ctx.surrogate_generate(1 << 24, seed=7)
        tr = ctx.surrogate_train(steps=60000, batch=16384, lr_max=2e-3, lr_min=2e-5, warmup=200,
                                 w_r=100.0, w_g=100.0, w_dir=1.0, seed=3)
        print("training", tr, flush=True)
        ctx.render_device()                               # warm
        t0 = time.perf_counter()
        ctx.render_device()
        plain_ms = (time.perf_counter() - t0) * 1e3
        p_fb = ctx.get_frame()
        pcls = ctx.classes()
        prgb = ctx.samples()
        ctx.render_surrogate(entropy_max=0.1, judge_fraction=0.02, lng_tol=0.005, device=True)   # warm
        t0 = time.perf_counter()
        _, st = ctx.render_surrogate(entropy_max=0.1, judge_fraction=0.02, lng_tol=0.005, device=True)
        sur_ms = (time.perf_counter() - t0) * 1e3
        s_fb = ctx.get_frame()
        scls = ctx.classes()
        srgb = ctx.samples()
        fbl, jl = ctx.surrogate_lists(st["fallback_rays"], st["judged_rays"])
    n = W * W * 4
    print(f"plain {plain_ms:.1f} ms, surrogate {sur_ms:.1f} ms ({plain_ms / sur_ms:.2f}x); {st}", flush=True)
    assert st["rays"] == n
    assert np.array_equal(scls[fbl], pcls[fbl])
    assert np.array_equal(srgb[fbl], prgb[fbl])
    agree = float(np.mean(scls == pcls))
    ja, jm = _judge_from_plain(jl, scls, pcls, srgb, prgb)
    ferr = _disk_pixel_flux_err(s_fb, p_fb, scls, pcls, 4)
    print(f"class agreement {agree:.5f}, judge {ja:.5f}/{jm:.4f}, disk-pixel flux err {ferr:.4f}, "
          f"fallback {st['fallback_rays'] / n:.4f}", flush=True)
    assert agree >= 0.995
    assert st["judge_class_agreement"] == pytest.approx(ja, abs=1e-12)
    assert st["judge_median_rel_err"] == pytest.approx(jm, rel=1e-5, abs=1e-7)
    assert 0.015 * n < st["judged_rays"] < 0.025 * n
    assert ferr <= 0.01
    assert st["fallback_rays"] < 0.05 * n
    assert sur_ms < 0.5 * plain_ms
