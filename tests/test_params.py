"""grrt_create's parameter resolution (reference src/api.cpp:43-126),
checked on the CPU through grrt_ext_resolve_params: the zero-means-default
rules, the Kerr spin fallback, Schwarzschild as Kerr a=0, the volumetric
disk's own wiring (raw mass and spin, r_out 30, derived peak temperature,
widened opacity band), and the horizon and ISCO closed forms."""
import ctypes as C
import math

import pytest


@pytest.fixture(scope="module")
def g():
    import grrt_amd
    grrt_amd.lib()
    return grrt_amd


def _zero(g):
    return g.GRRTParams()


def _isco(a):
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def test_all_zero_kerr_takes_every_default(g):
    p = _zero(g)
    p.width, p.height = 8, 6
    p.metric_type = g.METRIC_KERR
    p.disk_enabled = 1
    r = g.resolve_params(p)
    assert (r["width"], r["height"]) == (8, 6)
    assert r["M"] == 1.0
    assert r["a"] == 0.998                      # spin <= 0 -> 0.998
    assert r["r_obs"] == 50.0 and r["theta_obs"] == 1.396 and r["phi_obs"] == 0.0
    assert r["fov"] == 1.047
    assert r["max_steps"] == 10000 and r["tol"] == 1e-8
    assert r["disk"] == 1 and r["volumetric"] == 0
    assert r["r_out"] == 20.0 and r["t_disk"] == 1e7
    assert r["r_isco"] == pytest.approx(1.2369706, abs=1e-6)   # BPT72 prograde, a = 0.998
    assert r["r_in"] == r["r_isco"]
    assert r["r_horizon"] == pytest.approx(1 + math.sqrt(1 - 0.998 ** 2), rel=1e-15)
    assert r["spp"] == 1 and r["stars"] == 0


@pytest.mark.parametrize("spin", [0.0, -0.5, -1e-300])
def test_kerr_non_positive_spin_becomes_0998(g, spin):
    p = _zero(g)
    p.width = p.height = 4
    p.metric_type = g.METRIC_KERR
    p.spin = spin
    p.mass = 2.0
    r = g.resolve_params(p)
    assert r["M"] == 2.0 and r["a"] == pytest.approx(0.998 * 2.0, rel=1e-15)   # dimensional a = a* M


def test_schwarzschild_is_kerr_with_zero_spin(g):
    p = _zero(g)
    p.width = p.height = 4
    p.metric_type = g.METRIC_SCHWARZSCHILD
    p.spin = 0.7
    p.disk_enabled = 1
    r = g.resolve_params(p)
    assert r["a"] == 0.0 and r["r_horizon"] == 2.0 and r["r_isco"] == pytest.approx(6.0, rel=1e-14)
    assert r["r_in"] == r["r_isco"]


def test_explicit_values_are_kept(g):
    p = g.make_params(width=32, height=16, spin=0.5, observer_r=30.0, observer_theta=1.1, observer_phi=0.4,
                      fov=0.8, disk_inner=7.0, disk_outer=15.0, disk_temperature=3e6, tolerance=1e-6,
                      max_steps=77, spp=9, stars=True)
    r = g.resolve_params(p)
    assert (r["a"], r["r_obs"], r["theta_obs"], r["phi_obs"], r["fov"]) == (0.5, 30.0, 1.1, 0.4, 0.8)
    assert (r["r_in"], r["r_out"], r["t_disk"], r["tol"], r["max_steps"]) == (7.0, 15.0, 3e6, 1e-6, 77)
    assert r["spp"] == 9 and r["stars"] == 1
    assert r["r_isco"] == pytest.approx(_isco(0.5), rel=1e-13)


def test_disk_off(g):
    p = g.make_params(width=4, height=4, disk=False)
    r = g.resolve_params(p)
    assert r["disk"] == 0 and r["volumetric"] == 0


def test_volumetric_disk_takes_raw_spin_and_mass_and_r_out_30(g):
    """src/api.cpp:122-126: VolumetricDisk(params->mass, params->spin,
    disk_outer > 0 ? disk_outer : 30, ...), even for the Schwarzschild
    metric, while the metric itself follows the Kerr rules."""
    p = _zero(g)
    p.width = p.height = 4
    p.metric_type = g.METRIC_SCHWARZSCHILD
    p.spin = 0.6
    p.mass = 1.0
    p.disk_enabled = 1
    p.disk_volumetric = 1
    r = g.resolve_params(p)
    assert r["volumetric"] == 1 and r["disk"] == 0
    assert r["a"] == 0.0                          # the metric
    assert r["vol_spin"] == 0.6 and r["vol_M"] == 1.0
    assert r["vol_r_outer"] == 30.0
    assert r["vol_t_peak"] == 1e7
    # the default opacity band without frequency bins
    assert (r["vol_nu_min"], r["vol_nu_max"]) == (1e14, 1e16)


def test_volumetric_peak_temperature_from_mass_and_eddington(g):
    """T = 5e7 K eta^(1/4) (M/M_sun)^(-1/4) f_Edd^(1/4) with eta = 1 - E_isco:
    for a = 0, E_isco = sqrt(8/9) exactly."""
    p = _zero(g)
    p.width = p.height = 4
    p.metric_type = g.METRIC_KERR
    p.spin = 0.0
    p.disk_enabled = p.disk_volumetric = 1
    p.mass_solar, p.eddington_fraction = 10.0, 0.1
    r = g.resolve_params(p)
    eta = 1 - math.sqrt(8 / 9)
    assert r["vol_t_peak"] == pytest.approx(5e7 * eta ** 0.25 * 10.0 ** -0.25 * 0.1 ** 0.25, rel=1e-12)
    p.disk_temperature = 2e6                     # an explicit temperature wins
    assert g.resolve_params(p)["vol_t_peak"] == 2e6


def test_volumetric_opacity_band_widened_around_the_bins(g):
    p = g.make_params(width=4, height=4, volumetric=True, freqs=[3e15, 1e14, 2e17, 5e16])
    r = g.resolve_params(p)
    assert r["vol_nu_min"] == pytest.approx(1e13, rel=1e-15)
    assert r["vol_nu_max"] == pytest.approx(2e18, rel=1e-15)
    assert r["vol_r_outer"] == 20.0               # make_params sets disk_outer = 20
