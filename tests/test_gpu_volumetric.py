"""GPU parity of the volumetric-disk tracer (RGB and N-bin spectral) against
the CPU oracle on the same disk tables.

The volumetric ray-march takes classic RK4 steps without error control and
branches on position (volume exits, opacity exits, step-size rules), so a
rounding-level change of the path changes the intensity far above rounding:
the oracle built with the reference's own release flags (-O3 -ffast-math)
differs from the strict build by a median of ~1e-5.  The floor is measured
directly: the oracle built with one-ulp noise in every component of the Kerr
right-hand side (oracle variant "ulp") against the strict build.  On the
rays whose step sequence (outer attempts and steps, volume entries,
ray-march steps, all in the records) equals the strict oracle's, the GPU
must stay within 2x (median) and 3x (90th percentile) of that floor, keep
the oracle's sequence on as many rays as the one-ulp variant does (within
2%), and agree on the class of >= 99.5% of rays.  Measured: the GPU and the
one-ulp variant keep the same share of sequences (0.960 / 0.961 RGB,
0.982 / 0.982 spectral) with the same medians (1.8e-6 / 1.6-2.1e-6 RGB), so
the GPU's difference from the oracle is rounding noise carried through the
ray-march, which the reference shows against itself.  Mixed path (FP32 radiative transfer with compensated
summation): the error against the all-FP64 oracle is measured per frequency
and held to a median of 1e-3.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREQS = [10 ** (14 + 4 * k / 31) for k in range(32)]


def _setup(grrt, oracle, W, H, spin=0.9, theta=75.0, freqs=None, turbulence=0.4):
    p = grrt.make_params(width=W, height=H, metric=1, spin=spin, observer_theta=math.radians(theta),
                         fov=math.radians(90), volumetric=True, disk_outer=20.0, turbulence=turbulence,
                         seed=42, octaves=2, alpha=0.1, freqs=freqs)
    # the tables grrt_create builds (opacity tables and vertical columns on
    # the GPU, deterministic; the opacities agree with the host build to
    # ~1e-12 and the columns equal the host's over the same opacities), so
    # the oracle and the tracer read identical disks
    T = grrt.volumetric_tables(p, device=True)
    V = oracle.vol_struct(T)
    sc = oracle.scene(width=W, height=H, spin_a=spin, theta_obs=math.radians(theta),
                      fov=math.radians(90), disk=False)
    return p, V, sc


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    sel = (np.abs(a) > 0) & (np.abs(b) > 0)
    return np.abs(a[sel] - b[sel]) / np.abs(b[sel])


def _median_rel(a, b):
    r = _rel(a, b)
    return float(np.median(r)) if r.size else 0.0


def _same_sequence(g, o):
    """Rays whose outer attempts and steps, volume entries and ray-march
    steps all equal the oracle's."""
    return ((g["attempts"] == o["attempts"]) & (g["steps"] == o["steps"]) & (g["n_cross"] == o["n_cross"]) &
            (g["pad"] == o["pad"]))


def _pct(a, b):
    r = _rel(a, b)
    return " ".join(f"p{q}={np.percentile(r, q):.1e}" for q in (10, 50, 90)) if r.size else "empty"


def _noise_floor_check(name, g_vals, o_vals, u_vals, g_seq, u_seq):
    """GPU against the strict oracle, held to the oracle's own rounding-noise
    floor (the one-ulp variant against the strict build) on the rays whose
    step sequence is unchanged: median <= 2x and p90 <= 3x the floor."""
    e_g = _rel(g_vals[g_seq], o_vals[g_seq])
    e_u = _rel(u_vals[u_seq], o_vals[u_seq])
    mg, mu = float(np.median(e_g)), float(np.median(e_u))
    pg, pu = float(np.percentile(e_g, 90)), float(np.percentile(e_u, 90))
    print(f"{name}: same sequence GPU median {mg:.1e} p90 {pg:.1e}; one-ulp floor median {mu:.1e} p90 {pu:.1e}",
          flush=True)
    assert mg <= 2.0 * max(mu, 1e-9) and pg <= 3.0 * max(pu, 1e-8), (name, mg, mu, pg, pu)


def test_volumetric_rgb_parity(grrt, oracle):
    print("case: volumetric RGB a=0.9 64x64 pixel centres", flush=True)
    p, V, sc = _setup(grrt, oracle, 64, 64)
    with grrt.Context(p) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.set_records(True)
        fb = ctx.render()
        rec = ctx.records()
    ids = np.arange(64 * 64)
    orec, _ = oracle.vol_trace_samples(sc, V, oracle.JITTER_CENTER, 1, 0, ids)
    urec, _ = oracle.vol_trace_samples(sc, V, oracle.JITTER_CENTER, 1, 0, ids, variant="ulp")
    assert np.mean(rec["cls"] == orec["cls"]) >= 0.995
    g_seq = _same_sequence(rec, orec) & (rec["cls"] == orec["cls"])
    u_seq = _same_sequence(urec, orec) & (urec["cls"] == orec["cls"])
    print(f"rays with the oracle's step sequence: GPU {g_seq.mean():.3f}, one-ulp oracle {u_seq.mean():.3f}",
          flush=True)
    assert g_seq.mean() >= u_seq.mean() - 0.02
    for c in range(3):
        _noise_floor_check(f"channel {c}", rec["color"][:, c], orec["color"][:, c], urec["color"][:, c], g_seq, u_seq)
    assert np.isfinite(fb).all() and (fb[..., :3] >= 0).all()
    assert np.any(rec["cls"] == grrt.CLASS_DISK)


def test_volumetric_spectral_parity_fp64(grrt, oracle):
    print("case: volumetric spectral 32 bins a=0.9 48x48 hash 1spp FP64", flush=True)
    p, V, sc = _setup(grrt, oracle, 48, 48, freqs=FREQS)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(FREQS)
        ctx.set_records(True)
        cube = ctx.render_spectral()
        rec = ctx.records()
    ids = np.arange(48 * 48)
    orec, ospec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 1, 0, ids, freqs=FREQS)
    urec, uspec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 1, 0, ids, freqs=FREQS, variant="ulp")
    assert np.all(cube >= 0) and np.isfinite(cube).all()
    assert np.mean(rec["cls"] == orec["cls"]) >= 0.995
    g_seq = _same_sequence(rec, orec) & (rec["cls"] == orec["cls"])
    u_seq = _same_sequence(urec, orec) & (urec["cls"] == orec["cls"])
    print(f"rays with the oracle's step sequence: GPU {g_seq.mean():.3f}, one-ulp oracle {u_seq.mean():.3f}",
          flush=True)
    assert g_seq.mean() >= u_seq.mean() - 0.02
    flat = cube.reshape(-1, 32)
    for k in range(32):
        _noise_floor_check(f"bin {k}", flat[:, k], ospec[:, k], uspec[:, k], g_seq, u_seq)
    # pixels lit on one side only are the class-boundary grazers
    ocube = ospec.reshape(48, 48, 32)
    lit = (cube.sum(-1) > 0) != (ocube.sum(-1) > 0)
    assert lit.mean() <= 0.005


def test_volumetric_mixed_precision_error(grrt, oracle):
    print("case: volumetric spectral 32 bins 48x48 mixed precision", flush=True)
    p, V, sc = _setup(grrt, oracle, 48, 48, freqs=FREQS)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(FREQS)
        ctx.set_precision(True)
        cube = ctx.render_spectral_f32()
    assert cube.dtype == np.float32
    _, ospec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 1, 0, np.arange(48 * 48), freqs=FREQS)
    ocube = ospec.reshape(48, 48, 32)
    errs = [_median_rel(cube[..., k], ocube[..., k]) for k in range(32)]
    print("per-frequency median relative error (FP32 RT vs FP64 oracle):",
          " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) <= 1e-3


def test_spectral_cube_512sq_32bins_full_size(grrt, oracle):
    """512x512x32 float32 cube (Kerr a=0.9, theta=75 deg) in the mixed
    form: size-independent properties over the whole cube, ray-by-ray
    agreement with the FP64 oracle on sampled pixels."""
    print("case: spectral cube 512^2 x 32 bins, mixed precision", flush=True)
    W = 512
    p, V, sc = _setup(grrt, oracle, W, W, freqs=FREQS)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(FREQS)
        ctx.set_precision(True)
        cube = ctx.render_spectral_f32()
        st = ctx.stats()
    assert cube.shape == (W, W, 32) and cube.dtype == np.float32
    assert np.isfinite(cube).all() and (cube >= 0).all()
    lit = cube.sum(-1) > 0
    assert 0.01 < lit.mean() < 0.9
    rng = np.random.default_rng(11)
    pix = np.sort(rng.choice(np.flatnonzero(lit), 512, replace=False))
    _, ospec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 1, 0, pix, freqs=FREQS, threads=0)
    got = cube.reshape(-1, 32)[pix]
    errs = [_median_rel(got[:, k], ospec[:, k]) for k in range(32)]
    print("cube per-frequency median rel err:", " ".join(f"{e:.1e}" for e in errs),
          f"trace {st.trace_ms:.1f} ms", flush=True)
    assert max(errs) <= 1e-3


def test_far_field_fp32_error_stated(grrt, oracle):
    """Precision mode 2 (FP32 transfer and FP32 far-field geodesic, the
    spectral cube's precision policy): FP32 outer steps only on outbound
    far-field segments (r > 15 M, p_r > 0, away from the disk), so the
    emission is accumulated on FP64 geodesics; the error against the all-FP64
    oracle is stated per frequency: per-pixel median <= 1e-4, and the
    image-summed spectrum, which a few bright rays that turn back inwards
    after an FP32 segment dominate, within 5%."""
    print("case: volumetric spectral 32 bins 48x48, FP32 far field", flush=True)
    p, V, sc = _setup(grrt, oracle, 48, 48, freqs=FREQS)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(FREQS)
        ctx.set_precision(2)
        cube = ctx.render_spectral_f32()
    _, ospec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 1, 0, np.arange(48 * 48), freqs=FREQS)
    ocube = ospec.reshape(48, 48, 32)
    assert np.isfinite(cube).all() and (cube >= 0).all()
    med = [_median_rel(cube[..., k], ocube[..., k]) for k in range(32)]
    tot = [abs(float(cube[..., k].sum()) - float(ocube[..., k].sum())) / float(ocube[..., k].sum()) for k in range(32)]
    print("per-pixel median rel err:", " ".join(f"{e:.1e}" for e in med), flush=True)
    print("image-summed rel err:", " ".join(f"{e:.1e}" for e in tot), flush=True)
    assert max(med) <= 1e-4
    assert max(tot) <= 0.05


def test_spectral_fits_stream_equals_cube(grrt, oracle, tmp_path):
    """grrt_render_spectral_to_fits_cb streams the same cube grrt_render_spectral
    returns (BITPIX -64, bit for bit), and the float32 variant the mixed cube."""
    print("case: spectral FITS stream 1024x40 x 8 bins (two bands of rows)", flush=True)
    freqs = FREQS[::4]
    W, H = 1024, 40
    p, _, _ = _setup(grrt, oracle, W, H, freqs=freqs)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(freqs)
        cube = ctx.render_spectral()
        seen = []
        path = str(tmp_path / "s.fits")
        assert ctx.render_spectral_to_fits(path, progress=seen.append) == 0
        ctx.set_precision(True)
        cube32 = ctx.render_spectral_f32()
        path32 = str(tmp_path / "s32.fits")
        assert ctx.render_spectral_to_fits(path32, f32=True) == 0
    assert seen and seen[-1] == 1.0
    for pth, ref, dt in ((path, cube, ">f8"), (path32, cube32, ">f4")):
        raw = open(pth, "rb").read()
        off = 0
        while not raw[off:off + 80].startswith(b"END "):
            off += 80
        hdr = (off + 80 + 2879) // 2880 * 2880
        n = W * H * len(freqs)
        data = np.frombuffer(raw, dtype=dt, count=n, offset=hdr).reshape(len(freqs), H, W)
        assert np.array_equal(data[:, ::-1, :].transpose(1, 2, 0), ref)


def test_grouped_bins_equal_lane_per_ray(grrt, oracle):
    """The spectral kernel holds up to 256 bins in registers, spread over a
    group of lanes per ray; longer lists run one lane per ray with the bins in
    device scratch.  A 257-bin list made only of the 32 cube frequencies, with
    the step-control bin N/2 on the same frequency, takes the second path and
    must give every copy of each frequency bit for bit the 32-bin grouped
    result (all-FP64 and FP32-transfer modes)."""
    print("case: grouped vs lane-per-ray bins, 40x40, 32 vs 257 bins", flush=True)
    W = 40
    p, _, _ = _setup(grrt, oracle, W, W, freqs=FREQS)
    idx = list(range(16, 32)) + list(range(32)) * 7 + list(range(16)) + [16]
    assert len(idx) == 257 and idx[257 // 2] == 32 // 2
    big = [FREQS[i] for i in idx]
    with grrt.Context(p) as ctx:
        for mode in (0, 1):
            ctx.set_precision(mode)
            ctx.set_frequency_bins(FREQS)
            a = ctx.render_spectral() if mode == 0 else ctx.render_spectral_f32()
            ctx.set_frequency_bins(big)
            b = ctx.render_spectral() if mode == 0 else ctx.render_spectral_f32()
            assert (a.sum(-1) > 0).mean() > 0.05
            for k, i in enumerate(idx):
                assert np.array_equal(b[..., k], a[..., i]), (mode, k, i)


def test_volumetric_tile_and_spp_groups(grrt, oracle):
    """Lane groups over windows and several samples per pixel: a tile of the
    volumetric RGB frame equals the same pixels of the full frame bit for bit
    (4 hash-stratified samples per pixel), and a 4-spp spectral cube agrees
    with the oracle's samples averaged per pixel."""
    print("case: volumetric RGB 40x36 4 spp tile vs frame; spectral 24x20 4 spp vs oracle", flush=True)
    kw = dict(metric=1, spin=0.9, observer_theta=math.radians(75.0), fov=math.radians(90), volumetric=True,
              disk_outer=20.0, turbulence=0.4, seed=42, octaves=2, alpha=0.1, spp=4)
    with grrt.Context(grrt.make_params(width=40, height=36, **kw)) as ctx:
        full = ctx.render()
        tile = ctx.render_tile(7, 5, 19, 23)
    assert np.any(full[..., :3] > 0)
    assert np.array_equal(tile, full[5:28, 7:26])

    W, H = 24, 20
    p = grrt.make_params(width=W, height=H, freqs=FREQS, **kw)
    V = oracle.vol_struct(grrt.volumetric_tables(p, device=True))
    sc = oracle.scene(width=W, height=H, spin_a=0.9, theta_obs=math.radians(75.0), fov=math.radians(90),
                      disk=False)
    with grrt.Context(p) as ctx:
        ctx.set_frequency_bins(FREQS)
        ctx.set_records(True)
        cube = ctx.render_spectral()
        rec = ctx.records()
    ids = np.arange(W * H * 4)
    orec, ospec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 4, 0, ids, freqs=FREQS)
    urec, uspec = oracle.vol_trace_samples(sc, V, oracle.JITTER_HASH, 4, 0, ids, freqs=FREQS, variant="ulp")
    ocube = ospec.reshape(H, W, 4, 32).mean(2)
    ucube = uspec.reshape(H, W, 4, 32).mean(2)
    assert np.any(ocube > 0)
    assert np.mean(rec["cls"] == orec["cls"]) >= 0.995
    # a pixel keeps the oracle's sequence when all four of its samples do
    g_seq = (_same_sequence(rec, orec) & (rec["cls"] == orec["cls"])).reshape(-1, 4).all(1)
    u_seq = (_same_sequence(urec, orec) & (urec["cls"] == orec["cls"])).reshape(-1, 4).all(1)
    print(f"pixels with the oracle's step sequences: GPU {g_seq.mean():.3f}, one-ulp oracle {u_seq.mean():.3f}",
          flush=True)
    assert g_seq.mean() >= u_seq.mean() - 0.02
    # 480 pixels give too few lit ones per bin for a stable median, so the
    # bins are pooled (pixel x bin) into one noise-floor check.
    g_all, o_all, u_all = (c.reshape(-1, 32) for c in (cube, ocube, ucube))
    _noise_floor_check("4-spp, all 32 bins pooled", g_all.ravel(), o_all.ravel(), u_all.ravel(),
                       np.repeat(g_seq, 32), np.repeat(u_seq, 32))


def test_volumetric_step_doubling_parity(grrt, oracle):
    """The step-doubling RK4 controller (src/rk4.cpp:33-80) in the
    volumetric tracer's outer loop against the oracle's, RGB 48x48."""
    print("case: volumetric RGB a=0.9 48x48 pixel centres, step-doubling RK4", flush=True)
    p, V, sc = _setup(grrt, oracle, 48, 48)
    sc.integrator = 1
    with grrt.Context(p) as ctx:
        ctx.set_sampling(grrt.SAMPLING_CENTER)
        ctx.set_integrator(1)
        ctx.set_records(True)
        ctx.render()
        rec = ctx.records()
    ids = np.arange(48 * 48)
    orec, _ = oracle.vol_trace_samples(sc, V, oracle.JITTER_CENTER, 1, 0, ids)
    urec, _ = oracle.vol_trace_samples(sc, V, oracle.JITTER_CENTER, 1, 0, ids, variant="ulp")
    assert np.mean(rec["cls"] == orec["cls"]) >= 0.995
    g_seq = _same_sequence(rec, orec) & (rec["cls"] == orec["cls"])
    u_seq = _same_sequence(urec, orec) & (urec["cls"] == orec["cls"])
    print(f"rays with the oracle's step sequence: GPU {g_seq.mean():.3f}, one-ulp oracle {u_seq.mean():.3f}",
          flush=True)
    assert g_seq.mean() >= u_seq.mean() - 0.02
    for c in range(3):
        _noise_floor_check(f"step-doubling channel {c}", rec["color"][:, c], orec["color"][:, c],
                           urec["color"][:, c], g_seq, u_seq)
    assert np.any(rec["cls"] == grrt.CLASS_DISK)


def test_device_opacity_tables_equal_host(grrt):
    """grrt_create computes the opacity tables (one lane per (rho, T) cell)
    and the vertical columns (one wavefront per column) on the GPU; the host
    builder runs the same point functions and column steps on host threads.
    The device exp/pow/log of the opacity point functions differ from the
    host's in the last bits, so the tables agree to rounding, and so do the
    radial and vertical structure built from them; the Saha passes' 2-cycles
    (partial ionisation) are stable, so no cell drifts."""
    print("case: opacity tables host vs GPU, spectral cube scene (60 frequencies) and default disk", flush=True)
    for p in (grrt.make_params(width=64, height=64, metric=1, spin=0.9, volumetric=True, disk_outer=20.0,
                               turbulence=0.4, seed=42, octaves=2, alpha=0.1, freqs=FREQS),
              grrt.make_params(metric=1, spin=0.998, volumetric=True, disk_outer=30.0)):
        host = grrt.volumetric_tables(p)
        dev = grrt.volumetric_tables(p, device=True)
        print(f"n_nu {int(host['n_nu'])}: opacity tables host {host['build_ms_opacity']:.1f} ms, "
              f"GPU {dev['build_ms_opacity']:.1f} ms", flush=True)
        for name in ("kappa_abs", "kappa_es", "kappa_ross", "mu"):
            np.testing.assert_allclose(dev[name], host[name], rtol=1e-12, atol=1e-300, err_msg=name)
        for name in ("H", "rho_mid", "T_eff", "z_max", "T_profile"):
            np.testing.assert_allclose(dev[name], host[name], rtol=1e-9, atol=1e-300, err_msg=name)
        # Density profiles are normalised to 1 at the midplane; the column's
        # thin top (down to 1e-15) carries rounding of the opacities through
        # the hydrostatic integration, so they agree to 1e-9 of the midplane.
        np.testing.assert_allclose(dev["rho_profile"], host["rho_profile"], rtol=0, atol=1e-9)
        assert dev["rho_scale"] == pytest.approx(host["rho_scale"], rel=1e-9)
        assert np.array_equal(dev["perm"], host["perm"])


def test_device_columns_equal_host_columns_bit_for_bit(grrt):
    """Over the same (GPU-built) opacity tables, the 500 vertical columns
    built on the GPU equal the host threads' columns bit for bit: z_max,
    the density and the temperature profiles (both builds run the column
    steps of vol_column.hpp with IEEE +, -, *, /, sqrt only and no fused
    operations, so they round alike)."""
    print("case: vertical columns GPU vs host over the GPU opacity tables", flush=True)
    for p in (grrt.make_params(width=64, height=64, metric=1, spin=0.9, volumetric=True, disk_outer=20.0,
                               turbulence=0.4, seed=42, octaves=2, alpha=0.1, freqs=FREQS),
              grrt.make_params(metric=1, spin=0.998, volumetric=True, disk_outer=30.0),
              grrt.make_params(metric=1, spin=0.0, volumetric=True, disk_outer=20.0, alpha=0.3)):
        dev = grrt.volumetric_tables(p, device=True)
        hcol = grrt.volumetric_tables(p, device=True, device_columns=False)
        for name in ("kappa_es", "kappa_ross", "mu"):
            assert np.array_equal(dev[name], hcol[name]), name      # the same GPU opacity tables
        for name in ("z_max", "rho_profile", "T_profile", "H", "rho_mid", "T_eff"):
            same = dev[name] == hcol[name]
            print(f"{name}: {np.sum(~same)} of {same.size} entries differ", flush=True)
            assert np.array_equal(dev[name], hcol[name]), name
        assert dev["rho_scale"] == hcol["rho_scale"]
        print(f"total set-up: GPU columns {dev['build_ms_total']:.2f} ms, host columns "
              f"{hcol['build_ms_total']:.2f} ms", flush=True)
