"""Host-side volumetric disk set-up (no GPU): the reference's own opacity
and volumetric unit checks (tests/test_opacity.cpp:16-179,
tests/test_volumetric.cpp:15-68) restated against libgrrt.so's builder, plus
invariants of the construction.  The exact table contents have no
closed form and the reference cannot be built here, so beyond these checks
the table values are "parity unpinned"."""
import math

import numpy as np
import pytest

MP = 1.672622e-24


@pytest.fixture(scope="module")
def g():
    import grrt_amd
    return grrt_amd


@pytest.fixture(scope="module")
def disk(g):
    # VolumetricDisk(1.0, 0.998, 30.0, 1e7) with default VolumetricParams
    p = g.make_params(metric=1, spin=0.998, volumetric=True, disk_outer=30.0)
    return g.volumetric_tables(p)


@pytest.fixture(scope="module")
def calm(g):
    p = g.make_params(metric=1, spin=0.998, volumetric=True, disk_outer=30.0, turbulence=0.0)
    return g.volumetric_tables(p)


def test_saha_fully_ionized(g):
    s = g.saha(1e-8, 1e7)
    assert s["n_e"] == pytest.approx(1e-8 * 1.7 / (2 * MP), rel=0.05)
    assert s["mu"] == pytest.approx(0.6, rel=0.05)
    assert s["n_Hminus"] / s["n_e"] < 1e-10


def test_saha_partially_ionized(g):
    s = g.saha(1e-7, 6000.0)
    assert s["n_e"] < 0.5 * 1e-7 * 1.7 / (2 * MP)
    assert s["mu"] >= 0.8
    assert s["n_Hminus"] > 0.0


def test_saha_neutral(g):
    assert g.saha(1e-6, 3000.0)["mu"] == pytest.approx(1.3, rel=0.15)


def test_saha_limits(g):
    """Charge-neutral limits: fully ionised mean molecular weight of the
    X=0.70, Y=0.28, Z=0.02 mixture, mu = 1 / (2X + 3Y/4 + Z/2) ~ 0.62, and
    mu -> 1 / (X + Y/4 + ...) ~ 1.3 when neutral."""
    mu_ion = 1.0 / (2 * 0.70 + 0.75 * 0.28 + 0.5 * 0.02)
    assert g.saha(1e-12, 1e8)["mu"] == pytest.approx(mu_ion, rel=0.01)
    assert g.saha(1e-6, 50.0)["mu"] == 1.3


def test_opacities_positive_and_ordered(g):
    assert 0.0 < g.opacity(6e14, 1e-8, 1e6)["alpha_ff"] < math.inf
    o = g.opacity(6e14, 1e-7, 6000.0)
    assert o["alpha_hminus"] > o["alpha_ff"]                 # H- dominates at 6000 K
    assert 0.0 <= g.opacity(6e14, 1e-8, 3e4)["alpha_bf"] < math.inf
    assert 0.0 < g.opacity(6e14, 1e-9, 5e4)["kappa_abs"] < math.inf
    assert 0.0 < g.opacity(6e14, 1e-8, 6000.0)["B_nu"] < math.inf
    assert g.opacity(6e14, 1e-8, 1e7)["kappa_es"] == pytest.approx(0.34, rel=0.05)


def test_planck_against_textbook(g):
    """B_nu = 2 h nu^3 / c^2 / (exp(h nu / k T) - 1) with CODATA constants."""
    from scipy import constants as K
    h, c, k = K.h * 1e7, K.c * 100, K.k * 1e7     # CGS
    for nu, T in [(6e14, 6000.0), (1e15, 3e4), (1e13, 1e7)]:
        B = 2 * h * nu ** 3 / c ** 2 / math.expm1(h * nu / (k * T))
        assert g.opacity(nu, 1e-8, T)["B_nu"] == pytest.approx(B, rel=2e-6)


def test_opacity_tables(disk, oracle):
    V = oracle.vol_struct(disk)
    assert disk["n_nu"] == 20
    assert oracle.vol_kappa_abs(V, 6e14, 1e-8, 1e6) > 0
    assert oracle.vol_kappa_es(V, 1e-8, 1e7) == pytest.approx(0.34, rel=0.10)
    # mu and Rosseland tables at the same point (bilinear in log rho, log T)
    n = 100
    lr = (math.log10(1e-8) + 18) / 12 * (n - 1)
    lt = (7 - math.log10(3000)) / (8 - math.log10(3000)) * (n - 1)
    i, j = int(lr), int(lt)
    fr, ft = lr - i, lt - j
    mu = disk["mu"].reshape(n, n)
    m = (mu[i, j] * (1 - fr) * (1 - ft) + mu[i, j + 1] * (1 - fr) * ft + mu[i + 1, j] * fr * (1 - ft)
         + mu[i + 1, j + 1] * fr * ft)
    assert m == pytest.approx(0.6, rel=0.10)
    assert np.all(disk["kappa_ross"] >= 0) and np.all(np.isfinite(disk["kappa_abs"]))


def test_disk_construction(disk, calm, oracle):
    V = oracle.vol_struct(calm)
    H = np.interp(10.0, np.linspace(disk["r_min"], disk["r_outer"], 500), disk["H"])
    assert 0 < H < math.inf
    rho = [oracle.vol_density(V, 10.0, z * H, 0.0) for z in (0.0, 1.0, 3.0)]
    T = [oracle.vol_temperature(V, 10.0, z * H) for z in (0.0, 1.0, 3.0)]
    assert rho[0] > rho[1] > rho[2] and T[0] > T[1] > T[2]
    # taper: 1 outside the ISCO, small at the horizon
    def taper(r):
        return 1.0 if r >= disk["r_isco"] else math.exp(-((disk["r_isco"] - r) / disk["taper_width"]) ** 2)
    assert taper(disk["r_isco"]) == 1.0 and taper(disk["r_horizon"]) < 0.1
    assert oracle.vol_density(V, 50.0, 0.0, 0.0) == 0.0
    assert oracle.vol_density(V, 10.0, 100.0, 0.0) == 0.0


def test_normalisation_gives_tau_mid(disk):
    """rho_scale makes the midplane-to-surface Rosseland+Thomson depth at
    the peak-density radius equal to tau_mid = 100 (src/volumetric_disk.cpp:783-849)."""
    r = np.linspace(disk["r_min"], disk["r_outer"], 500)
    rm = np.where(r >= disk["r_isco"], disk["rho_mid"], 0.0)
    k = int(np.argmax(rm))
    col = disk["rho_profile"].reshape(500, 128)[k]
    dz = disk["z_max"][k] / 127
    integral = 2 * np.sum(0.5 * (col[1:] + col[:-1]) * dz)
    assert integral > 0 and disk["rho_scale"] > 0
    assert disk["noise_scale"] == pytest.approx(max(2 * disk["H"][k], 0.01))
    # kappa(rho_mid_cgs, T_eff) * rho_scale * rho_mid * column = tau_mid, with
    # kappa = Rosseland + Thomson from the tables (self-consistent to ~1%).
    def lookup(tab, rho, T):
        n = 100
        lr = min(max((math.log10(rho) + 18) / 12 * (n - 1), 0), n - 2)
        lt = min(max((math.log10(T) - math.log10(3000)) / (8 - math.log10(3000)) * (n - 1), 0), n - 2)
        i, j = int(lr), int(lt)
        fr, ft = lr - i, lt - j
        t = tab.reshape(n, n)
        return (t[i, j] * (1 - fr) * (1 - ft) + t[i, j + 1] * (1 - fr) * ft + t[i + 1, j] * fr * (1 - ft)
                + t[i + 1, j + 1] * fr * ft)
    rho_c = min(max(disk["rho_scale"] * rm[k], 1e-18), 1e-6)
    Tc = min(max(disk["T_eff"][k], 3000.0), 1e8)
    kap = lookup(disk["kappa_ross"], rho_c, Tc) + lookup(disk["kappa_es"], rho_c, Tc)
    assert kap * disk["rho_scale"] * rm[k] * integral == pytest.approx(100.0, rel=0.02)


def test_vertical_profiles_physical(disk):
    prof = disk["rho_profile"].reshape(500, 128)
    T = disk["T_profile"].reshape(500, 128)
    assert np.all(prof[:, 0] == 1.0)                 # normalised to the midplane
    assert np.all(np.diff(T[prof[:, 1] > 0], axis=1) <= 1e-6 * T[prof[:, 1] > 0][:, :1])  # T falls with |z|
    zr = disk["z_max"] / disk["H"]
    assert np.all((zr >= 3.0 - 1e-12) & (zr <= 20.0 + 1e-12))


def _mt_raw(seed, n):
    rs = np.random.RandomState(seed)
    return rs.randint(0, 2 ** 32, size=n, dtype=np.uint64)


def test_noise_permutation_matches_independent_mt19937(disk):
    """Fisher–Yates from 255 down with uniform_int_distribution(0, i) on
    mt19937(42): libstdc++ draws one 32-bit word and maps it by Lemire's
    multiply-and-reject method (an independent MT19937 from numpy)."""
    raw = iter(_mt_raw(42, 4096).tolist())
    base = list(range(256))
    for i in range(255, 0, -1):
        rng = i + 1
        while True:
            x = next(raw)
            m = x * rng
            low = m & 0xFFFFFFFF
            if low < rng:
                thresh = (2 ** 32 - rng) % rng
                if low < thresh:
                    continue
            break
        j = m >> 32
        base[i], base[j] = base[j], base[i]
    perm = disk["perm"]
    assert sorted(perm[:256].tolist()) == list(range(256))
    assert np.array_equal(perm[:256], perm[256:])
    assert perm[:256].tolist() == base
