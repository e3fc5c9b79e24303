"""The volumetric disk's radial flux and vertical density profiles
(src/volumetric_disk.cpp:279-346 and :505-777, restated in
grrt_amd/csrc/volumetric_host.cpp), pinned against independent solutions on
the CPU:

* the vertical density of each column steps, node by node, by classic RK4
  (stage values floored at 1e-15) through hydrostatic equilibrium with gas
  and flux-limited radiation pressure,
      d(rho c_s^2 + f E)/dz = -rho Omega_z^2 z   (z in M, c_s^2 / c^2),
  over the column's final temperature profile, with the coefficients of
  the pass's uniform start column (mu, Levermore-Pomraning f, E = a T^4)
  and central differences on the 128 nodes.  The column build takes log10
  for its table lookups from its own series (host and GPU builds alike),
  pinned here against math.log10 to 4e-16.  The columns of this disk are
  radiation dominated: the density leaves the midplane downwards at once
  and the profile is the floored map's realisation, which is rounding
  sensitive; a one-step check is not.  Dropping the radiation term, taking
  the orbital for the vertical epicyclic frequency or an Euler for the RK4
  step fails it by orders of magnitude (the d(c_s^2)/dz term is at rounding
  level along these isothermal columns);
* the radial flux integral (the reference's integrand over [r_isco, r]) at
  every radial node, against adaptive quadrature with analytic derivatives
  of the circular-orbit E(r), L(r), Omega(r);
* the radial scale height and midplane density, against the thin alpha-disk
  relations H = c_s / Omega_z and local dissipation balance
  (nu Sigma (r dOmega/dr)^2 proportional to T_eff^4).
No GPU is needed: the tables come from the library's host build."""
import math

import numpy as np
import pytest

scipy_integrate = pytest.importorskip("scipy.integrate")

K_B, M_P, C_LIGHT, SIGMA_SB = 1.380649e-16, 1.672622e-24, 2.997924e10, 5.670374e-5
A_RAD = 4.0 * SIGMA_SB / C_LIGHT
TAU_MID = 100.0
LOG_RHO = (-18.0, -6.0)
LOG_T = (math.log10(3000.0), 8.0)


@pytest.fixture(scope="module")
def disk():
    import grrt_amd as g
    p = g.make_params(width=64, height=64, metric=g.METRIC_KERR, spin=0.9, observer_theta=math.radians(75.0),
                      volumetric=True, disk_outer=20.0, turbulence=0.4, seed=42, octaves=2, alpha=0.1,
                      freqs=[10 ** (14 + 4 * k / 31) for k in range(32)])
    return g.volumetric_tables(p)


def _log10(x):
    """The column build's log10: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
    ln m = 2 atanh((m - 1)/(m + 1)) by its series through s^21, then
    (e ln 2 + ln m) / ln 10, in the library's operation order."""
    m, e = math.frexp(x)
    if m < 0.70710678118654752440:
        m, e = m * 2.0, e - 1
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    p = 1.0 / 21.0
    for k in (19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = p * s2 + 1.0 / k
    lnm = 2.0 * s + 2.0 * s * (s2 * p)
    de = float(e)
    return ((de * 6.93147180369123816490e-01 + lnm) + de * 1.90821492927058770002e-10) * 0.43429448190325182765


def test_column_log10_is_log10():
    xs = np.concatenate([10.0 ** np.linspace(-18, 8, 20001), np.random.default_rng(1).uniform(1e-18, 1e8, 2000)])
    err = max(abs(_log10(float(x)) - math.log10(x)) / max(abs(math.log10(x)), 1e-300) for x in xs
              if abs(math.log10(x)) > 1e-3)
    assert err < 4e-16
    assert abs(_log10(1.0)) == 0.0 and abs(_log10(1000.0) - 3.0) < 1e-15


def _bilinear(lut, rho, T, n=100):
    """Bilinear interpolation in (log10 rho, log10 T) over the 100 x 100
    opacity grid (rho-major), arguments clamped to the grid; scalar, in the
    library's operation order."""
    rho = min(max(rho, 1e-18), 1e-6)
    T = min(max(T, 3000.0), 1e8)

    def cell(v, lo, hi):
        t = (v - lo) / (hi - lo) * (n - 1)
        t = min(max(t, 0.0), float(n - 2))
        i = int(t)
        return i, t - i
    ir, fr = cell(_log10(rho), math.log10(1e-18), math.log10(1e-6))
    it, ft = cell(_log10(T), math.log10(3000.0), math.log10(1e8))
    return (lut[ir * n + it] * (1 - fr) * (1 - ft) + lut[ir * n + it + 1] * (1 - fr) * ft +
            lut[(ir + 1) * n + it] * fr * (1 - ft) + lut[(ir + 1) * n + it + 1] * fr * ft)


def _column_map(D, i, radiation=True, gas_gradient=True, vertical_frequency=True):
    """Column i's coefficients and its right-hand side.  The stored profile
    is the last pass's; when that pass still ended above 1e-10 at the top
    (and below 20 H), the column top was raised by H after it, so the
    profile's node spacing is (z_max - H) / 127.

    Along these columns T is constant to ~1e-7, so the gradients of c_s^2
    and f E are differences of nearly equal numbers: each product and
    difference is formed here in the library's order (IEEE double, no
    fused operations), which makes them equal to the last bit."""
    NR, NZ = int(D["n_r"]), int(D["n_z"])
    r = D["r_min"] + (D["r_outer"] - D["r_min"]) * i / (NR - 1)
    H, Te, zmax = float(D["H"][i]), float(D["T_eff"][i]), float(D["z_max"][i])
    lib = D["rho_profile"].reshape(NR, NZ)[i]
    if lib[-1] > 1e-10 and zmax - H < 20.0 * H:
        zmax = zmax - H
    T = [float(v) for v in D["T_profile"].reshape(NR, NZ)[i]]
    a, M = float(D["a"]), float(D["M"])
    kr, ke, km = D["kappa_ross"], D["kappa_es"], D["mu"]

    def omega_z2(rr):
        # vertical epicyclic frequency of circular Kerr orbits (Okazaki, Kato & Fukue 1987)
        om = math.sqrt(M) / (rr * math.sqrt(rr) + a * math.sqrt(M))
        if not vertical_frequency:              # a planted mistake: the orbital frequency
            return om * om
        return om * om * (1.0 - 4.0 * a * math.sqrt(M) / math.sqrt(rr * rr * rr) + 3.0 * a * a / (rr * rr))
    Omz2 = omega_z2(r)
    if r < D["r_isco"] or Omz2 <= 0:
        Omz2 = omega_z2(float(D["r_isco"]))
    kref = max(_bilinear(kr, 1e-10, Te) + _bilinear(ke, 1e-10, Te), 1.0)
    rho_ref = min(max(TAU_MID / (kref * 3.0 * H), 1e-18), 1e-6)
    dz = zmax / (NZ - 1)
    # the pass's coefficients come from its uniform start column (rho_hat = 1)
    E = [A_RAD * t * t * t * t for t in T]
    mu = []
    for t in T:
        m = _bilinear(km, rho_ref, t)
        mu.append(m if (m > 0 and math.isfinite(m)) else 0.6)
    f = []
    for z in range(NZ):
        dE = 0.0
        if z == NZ - 1:
            dE = (E[z] - E[z - 1]) / dz
        elif z > 0:
            dE = (E[z + 1] - E[z - 1]) / (2.0 * dz)
        den = _bilinear(kr, rho_ref, T[z]) * rho_ref * E[z]
        R = 1e30 if den < 1e-30 else abs(dE) / den
        lam = (2.0 + R) / (6.0 + 3.0 * R + R * R)      # Levermore & Pomraning (1981)
        f.append(lam + lam * lam * R * R)
    cs2 = [K_B * T[z] / (mu[z] * M_P) for z in range(NZ)]
    dcs2, dfE = [0.0] * NZ, [0.0] * NZ
    for z in range(1, NZ):
        if z == NZ - 1:
            dcs2[z] = (cs2[z] - cs2[z - 1]) / dz
            dfE[z] = (f[z] * E[z] - f[z - 1] * E[z - 1]) / dz
        else:
            dcs2[z] = (cs2[z + 1] - cs2[z - 1]) / (2.0 * dz)
            dfE[z] = (f[z + 1] * E[z + 1] - f[z - 1] * E[z - 1]) / (2.0 * dz)
    c2 = C_LIGHT * C_LIGHT

    def rhs(ze, y):
        zf = ze / dz
        k = min(max(int(zf), 0), NZ - 2)
        t = zf - k
        cs = K_B * ((1.0 - t) * T[k] + t * T[k + 1]) / (((1.0 - t) * mu[k] + t * mu[k + 1]) * M_P)
        dc = (1.0 - t) * dcs2[k] + t * dcs2[k + 1]
        df = (1.0 - t) * dfE[k] + t * dfE[k + 1]
        if cs < 1e-30:
            return 0.0
        # d(rho c_s^2)/dz + d(fE)/dz = -rho Omega_z^2 z, everything over c^2
        num = -y * (Omz2 * ze)
        if gas_gradient:
            num = num - y * (dc / c2)
        if radiation:
            num = num - df / (rho_ref * c2)
        return num / max(cs / c2, 1e-30)
    return dz, rhs


def _one_step(rhs, z, y, dz, floor=1e-15, euler=False):
    """One classic RK4 step with the stage values kept above the floor
    (euler: a planted mistake, one explicit Euler step)."""
    k1 = dz * rhs(z, y)
    if euler:
        return max(y + k1, floor)
    k2 = dz * rhs(z + 0.5 * dz, max(y + 0.5 * k1, floor))
    k3 = dz * rhs(z + 0.5 * dz, max(y + 0.5 * k2, floor))
    k4 = dz * rhs(z + dz, max(y + k3, floor))
    return max(y + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0, floor)


def _step_errors(D, i, euler=False, **kw):
    NR, NZ = int(D["n_r"]), int(D["n_z"])
    lib = D["rho_profile"].reshape(NR, NZ)[i]
    dz, rhs = _column_map(D, i, **kw)
    pred = np.array([_one_step(rhs, k * dz, float(lib[k]), dz, euler=euler) for k in range(NZ - 1)])
    return np.abs(pred - lib[1:]) / np.maximum(np.abs(lib[1:]), 1e-15)


COLUMNS = (0, 30, 60, 100, 160, 250, 290, 350, 450, 498)


def test_vertical_density_steps_solve_hydrostatic_equilibrium(disk):
    """Every node of every sampled column follows from the node below it by
    one RK4 step of hydrostatic equilibrium with the floor.  These columns
    are radiation dominated: in exact arithmetic the density leaves the
    midplane downwards at once, and the stored profile is the floored map's
    realisation, so the check is node by node, not against a continuous
    solution."""
    D = disk
    assert D["rho_profile"].reshape(int(D["n_r"]), -1)[:, 0].min() == 1.0
    worst = 0.0
    for i in COLUMNS:
        assert D["H"][i] > 0 and D["T_eff"][i] > 0 and D["rho_mid"][i] > 0, i
        err = _step_errors(D, i)
        worst = max(worst, float(err.max()))
        assert err.max() < 1e-9, (i, int(err.argmax()), float(err.max()))
    print(f"worst one-step relative difference over {len(COLUMNS)} columns: {worst:.2e}")


def test_vertical_check_catches_planted_mistakes(disk):
    """The same node-by-node check misses by far more than the bar with a
    planted mistake: the radiation term left out, the orbital frequency in
    place of the vertical epicyclic one, an Euler step in place of RK4.
    (Along these columns T is constant to ~1e-7, so the gas-pressure
    gradient term is at rounding level: leaving it out moves the profile by
    ~1e-7, and a mistake in that term alone would pass.)"""
    D = disk
    for kw in (dict(radiation=False), dict(vertical_frequency=False), dict(euler=True)):
        worst = max(float(_step_errors(D, i, **kw).max()) for i in COLUMNS)
        assert worst > 1e-3, (kw, worst)
    gas = max(float(_step_errors(D, i, gas_gradient=False).max()) for i in COLUMNS)
    assert gas < 1e-6


def test_columns_without_flux(disk):
    """Where the flux table gives no T_eff (the outer edge), the column is
    the placeholder of src/volumetric_disk.cpp:505-777: z_max = 3 H, the
    density 1 at the midplane and 0 above, T = T_eff = 0."""
    D = disk
    NR, NZ = int(D["n_r"]), int(D["n_z"])
    cold = np.where(D["T_eff"] <= 0)[0]
    assert len(cold) >= 1 and cold[-1] == NR - 1
    for i in cold:
        rp = D["rho_profile"].reshape(NR, NZ)[i]
        assert rp[0] == 1.0 and np.all(rp[1:] == 0.0)
        assert D["z_max"][i] == 3.0 * D["H"][i]
        assert np.all(D["T_profile"].reshape(NR, NZ)[i] == 0.0)


def test_column_top_rule(disk):
    """z_max is 3 H plus whole H extensions while the top stays above 1e-10
    of the midplane, at most 20 H and at most eight passes."""
    D = disk
    k = D["z_max"] / D["H"]
    ok = D["H"] > 0
    assert np.allclose(k[ok], np.round(k[ok]), atol=1e-9)
    assert k[ok].min() >= 3 - 1e-9 and k[ok].max() <= 11 + 1e-9


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_radial_flux_integral_against_quadrature(disk):
    """T_eff at every radial node is t_peak (F / F_max)^(1/4) with
    F(r) = 3 M / (8 pi r^3) (E - Omega L)^-1 (-dOmega/dr)
           * int_{r_isco}^{r} [(E - E_isco) L' - (L - L_isco) E'] dr'
    (src/volumetric_disk.cpp:279-346), here with analytic derivatives and
    adaptive quadrature; the library's trapezoid over the 500 nodes with
    central differences agrees to 3e-4 in T_eff beyond r_isco + M (median
    over the disk 2e-5)."""
    D = disk
    M, a, ri, ro = D["M"], D["a"], D["r_isco"], D["r_outer"]
    N = int(D["n_r"])

    def wk(r):
        return math.sqrt(M / r ** 3)

    def E(r):
        aw = a * wk(r)
        return (1 - 2 * M / r + aw) / math.sqrt(1 - 3 * M / r + 2 * aw)

    def L(r):
        aw = a * wk(r)
        return math.sqrt(M * r) * (1 - 2 * aw + a * a / (r * r)) / math.sqrt(1 - 3 * M / r + 2 * aw)

    def Om(r):
        return wk(r) / (1 + a * wk(r))

    def d(fn, r, h=1e-5):           # fourth-order central difference, well inside the smooth region
        return (-fn(r + 2 * h) + 8 * fn(r + h) - 8 * fn(r - h) + fn(r - 2 * h)) / (12 * h)

    Ei, Li = E(ri), L(ri)

    def integrand(r):
        return (E(r) - Ei) * d(L, r) - (L(r) - Li) * d(E, r)

    rs = ri + (ro - ri) * np.arange(1, N) / (N - 1)
    F = np.zeros(N)
    acc, prev_r = 0.0, ri
    for k, r in enumerate(rs, start=1):
        acc += scipy_integrate.quad(integrand, prev_r, r, epsabs=0, epsrel=1e-12)[0]
        prev_r = r
        F[k] = max(3 * M / (8 * math.pi * r ** 3) / (E(r) - Om(r) * L(r)) * (-d(Om, r)) * acc, 0.0)
    Fmax = F.max()
    # T_eff on the library's radial grid (which starts at r_min, not r_isco)
    grid = D["r_min"] + (ro - D["r_min"]) * np.arange(N) / (N - 1)
    Fr = np.interp(grid, ri + (ro - ri) * np.arange(N) / (N - 1), F)
    sel = (grid > ri + 0.05) & (grid < ro - 1e-9)
    Te = D["t_peak"] * (Fr[sel] / Fmax) ** 0.25
    err = np.abs(D["T_eff"][sel] - Te) / Te
    far = grid[sel] > ri + 1.0
    print(f"T_eff: max relative difference {err.max():.2e} (beyond r_isco + M: {err[far].max():.2e}), "
          f"median {np.median(err):.2e}")
    # the trapezoid's first intervals carry a large relative error where
    # the flux starts from zero at the ISCO; beyond r_isco + M it is small
    assert err[far].max() < 3e-4 and np.median(err) < 1e-4


def test_scale_height_and_midplane_density_are_the_alpha_disk_relations(disk):
    """The radial structure (src/volumetric_disk.cpp:348-483) against the
    thin alpha-disk relations it implements:
      * H = c_s / Omega_z, the isothermal sound speed c_s^2 = k T_eff /
        (mu m_p) over the vertical epicyclic frequency of circular Kerr
        orbits, Omega_z^2 = Omega^2 (1 - 4 a r^-3/2 + 3 a^2 r^-2) with
        Omega = 1 / (r^3/2 + a) (Okazaki, Kato & Fukue 1987; M = 1);
      * local dissipation balance: the viscous heating nu Sigma (r dOmega/dr)^2,
        with nu = alpha c_s H and Sigma = sqrt(2 pi) H rho_mid, equals the
        emitted flux F = sigma T_eff^4 up to one constant (the disk's F_max
        over t_peak^4), so rho_mid H^2 c_s (r dOmega/dr)^2 / T_eff^4 is the
        same at every radius.
    Over radii outside the ISCO where H is not clamped to [0.001, 5] M.  The
    disk is fully ionised there (T_eff > 1e6 K), so mu is the Saha value at
    any density.  Taking the orbital for the vertical frequency, mu = 1, or
    Sigma = 2 H rho_mid with rho_mid proportional to 1/H each breaks one of
    the two by far more than the bars."""
    import grrt_amd as g
    from scipy import constants as K
    D = disk
    M, a, N = D["M"], D["a"], int(D["n_r"])
    assert M == 1.0
    r = D["r_min"] + (D["r_outer"] - D["r_min"]) * np.arange(N) / (N - 1)
    T, H, rho = D["T_eff"], D["H"], D["rho_mid"]
    sel = (r > D["r_isco"] + 0.05) & (T > 2e5) & (rho > 0.0) & (H > 0.001 * 1.001) & (H < 5.0 * 0.999)
    assert sel.sum() > 300
    r, T, H, rho = r[sel], T[sel], H[sel], rho[sel]
    om = 1.0 / (r ** 1.5 + a)
    omz = om * np.sqrt(1.0 - 4.0 * a / r ** 1.5 + 3.0 * a * a / (r * r))
    mu = np.array([g.saha(1e-10, t)["mu"] for t in T])
    assert np.ptp(mu) / mu.mean() < 1e-6                       # fully ionised
    cs = np.sqrt(K.k * 1e7 * T / (mu * K.m_p * 1e3)) / (K.c * 1e2)
    assert np.abs(H / (cs / omz) - 1.0).max() < 1e-5
    dom = -1.5 * np.sqrt(r) / (r ** 1.5 + a) ** 2               # d Omega / dr
    q = rho * H * H * cs * (r * dom) ** 2 / T ** 4
    assert q.max() / q.min() - 1.0 < 1e-9, (q.min(), q.max())
