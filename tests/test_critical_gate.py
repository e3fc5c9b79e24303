"""The critical-curve score behind the surrogate's fallback gate
(grrt_ext_surrogate_set_critical_gate), restated in numpy in
tests/parity_util.py, pinned by the cases the mathematics fixes: the
Schwarzschild critical impact parameter 3 sqrt(3) M, the Kerr spherical
photon orbits (Bardeen 1973; Teo 2003), and its sign against the oracle
integrator's outcome classes."""
import math

import numpy as np

from parity_util import critical_score, photon_orbit_radii


def test_photon_orbit_radii_known_values():
    assert photon_orbit_radii(0.0) == (3.0, 3.0)
    r0, r1 = photon_orbit_radii(1.0)
    assert abs(r0 - 1.0) < 1e-12 and abs(r1 - 4.0) < 1e-12
    r0, r1 = photon_orbit_radii(0.998)
    assert 1.07 < r0 < 1.08 and 3.99 < r1 < 4.0


def test_schwarzschild_critical_impact_parameter():
    # a = 0: critical at b^2 + q = 27 M^2, for any split between b and q
    for theta in (0.3, 1.0, 1.39626):
        for b in (0.0, 2.0, 4.0, 5.19):
            q = 27.0 - b * b
            s, c = math.sin(theta), math.cos(theta)
            pth2 = q - c * c * b * b / (s * s)
            if pth2 < 0:
                continue
            pt, pth, pph = -1.0, math.sqrt(pth2), b
            assert abs(critical_score(0.0, 1.0, pt, pth, pph, theta)[0]) < 1e-12
            # larger impact parameter: a turning point outside r = 3 (score < 0)
            assert critical_score(0.0, 1.0, pt, pth * 1.01, pph * 1.01, theta)[0] < 0
            assert critical_score(0.0, 1.0, pt, pth * 0.99, pph * 0.99, theta)[0] > 0


def test_kerr_spherical_photon_orbits_score_zero():
    # (xi, eta) of the spherical photon orbit at radius r (M = 1):
    #   xi  = -(r^3 - 3 r^2 + a^2 r + a^2) / (a (r - 1))
    #   eta = r^3 (4 a^2 - r (r - 3)^2) / (a^2 (r - 1)^2)
    for a in (0.5, 0.9, 0.998):
        r0, r1 = photon_orbit_radii(a)
        for f in (0.05, 0.3, 0.5, 0.7, 0.95):
            r = r0 + f * (r1 - r0)
            xi = -(r ** 3 - 3 * r * r + a * a * r + a * a) / (a * (r - 1))
            eta = r ** 3 * (4 * a * a - r * (r - 3) ** 2) / (a * a * (r - 1) ** 2)
            assert eta >= -1e-9
            theta = math.pi / 2          # at the equator p_theta^2 = eta
            pt, pth, pph = -1.0, math.sqrt(max(eta, 0.0)), xi
            sc = critical_score(a, 1.0, pt, pth, pph, theta)[0]
            assert abs(sc) < 1e-5, (a, r, sc)
            # one percent off the orbit's impact parameters is far from zero
            assert abs(critical_score(a, 1.0, pt, pth * 1.01, pph * 1.01, theta)[0]) > 1e-3


def test_score_sign_matches_the_integrator(oracle):
    """Rays of a disk-free Kerr a=0.998 view zoomed on the shadow: a score
    above +0.02 means capture, below -0.02 escape, per the oracle's own
    FP64 integration."""
    a = 0.998
    sc = oracle.scene(width=48, height=48, spin_a=a, theta_obs=math.radians(80.0), fov=math.radians(25.0),
                      disk=False)
    ids = np.arange(48 * 48)
    st = np.array([oracle.camera_ray(sc, (k % 48) + 0.5, (k // 48) + 0.5) for k in ids])
    s = critical_score(a, 1.0, st[:, 4], st[:, 6], st[:, 7], st[:, 2])
    rays = oracle.trace_samples(sc, oracle.JITTER_CENTER, 1, 0, ids, threads=8)
    cls = rays["cls"]
    firm = np.abs(s) > 0.02
    assert firm.sum() > 1500 and (s > 0.02).sum() > 300 and (s < -0.02).sum() > 300
    assert np.all(cls[s > 0.02] == 0)      # captured
    assert np.all(cls[s < -0.02] == 1)     # escaped
    print(f"near-critical (|s| <= 0.02): {np.sum(~firm)} of {len(s)}")
