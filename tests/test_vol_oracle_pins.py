"""Pins of the oracle's volumetric runtime by what the mathematics fixes (no
GPU).  The oracle's ray-march (oracle/path_oracle.cpp) is the checker of the
GPU volumetric path; these checks keep a plausible mistake in it from
passing unnoticed:

- simplex noise (src/noise.cpp:38-115) against an independent NumPy
  restatement of Gustavson's published 3D simplex noise (S. Gustavson,
  "Simplex noise demystified", 2005) with the disk's permutation, and the
  fBm density modulation (src/volumetric_disk.cpp:226-257) against it;
- density and temperature (src/volumetric_disk.cpp:164-273) against the
  tables at grid nodes, linear interpolation at midpoints, the ISCO taper,
  the Gaussian edge above z_max and the zero regions;
- the gas 4-velocity (src/volumetric_disk.cpp:95-133): unit norm,
  Keplerian angular velocity outside the ISCO, conserved E_isco, L_isco
  inside it, both against the Bardeen-Press-Teukolsky closed forms;
- the opacity lookups (src/opacity.cpp:307-346): table values at nodes,
  log-trilinear / bilinear at cell midpoints, and the reference's index
  clamp at the top node;
- the backward transfer step (src/geodesic_tracer.cpp:311-315, :724-728)
  against the formal solution of the transfer equation for piecewise
  constant sources, and a disk without matter against zero intensity;
- the radial T_eff table (src/volumetric_disk.cpp:279-400) against
  adaptive quadrature of the Novikov-Thorne flux.
"""
import math

import numpy as np
import pytest

M = 1.0


@pytest.fixture(scope="module")
def g():
    import grrt_amd
    grrt_amd.lib()
    return grrt_amd


@pytest.fixture(scope="module")
def params(g):
    # the spectral-cube disk: Kerr a=0.9, r_out 20, alpha 0.1, turbulence 0.4, 2 octaves, seed 42
    freqs = [10 ** (14 + 4 * k / 31) for k in range(32)]
    return g.make_params(width=64, height=64, metric=1, spin=0.9, observer_theta=math.radians(75), volumetric=True,
                         disk_outer=20.0, turbulence=0.4, seed=42, octaves=2, alpha=0.1, freqs=freqs)


@pytest.fixture(scope="module")
def disk(g, params):
    return g.volumetric_tables(params)


@pytest.fixture(scope="module")
def V(oracle, disk):
    return oracle.vol_struct(disk)


@pytest.fixture(scope="module")
def Vcalm(oracle, disk):
    return oracle.vol_struct({**disk, "turbulence": 0.0})


# ---- simplex noise -------------------------------------------------------

GRAD3 = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                  [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], float)


def simplex3(perm, x, y, z):
    """Gustavson's 3D simplex noise, vectorised: skew to the cubic lattice,
    find the simplex by sorting the offsets, sum the four corners' radially
    attenuated gradient ramps (0.6 - d^2)^4 (g . d), scale by 32."""
    perm = np.asarray(perm, np.int64)
    P = np.stack([x, y, z], -1).astype(np.float64)
    s = P.sum(-1, keepdims=True) / 3.0
    cell = np.floor(P + s)
    t = cell.sum(-1, keepdims=True) / 6.0
    d0 = P - (cell - t)
    # corner offsets in lattice units, largest offset component first
    order = np.argsort(-d0, axis=-1, kind="stable")
    e = np.eye(3, dtype=np.int64)
    c1 = e[order[:, 0]]
    c2 = c1 + e[order[:, 1]]
    corners = [np.zeros_like(c1), c1, c2, np.ones_like(c1)]
    ijk = cell.astype(np.int64) & 255
    total = np.zeros(len(P))
    for n, c in enumerate(corners):
        d = d0 - c + n / 6.0
        h = perm[ijk[:, 0] + c[:, 0] + perm[ijk[:, 1] + c[:, 1] + perm[ijk[:, 2] + c[:, 2]]]] % 12
        w = 0.6 - (d * d).sum(-1)
        total += np.where(w > 0, w ** 4 * (GRAD3[h] * d).sum(-1), 0.0)
    return 32.0 * total


def test_simplex_noise_is_gustavsons(oracle, disk):
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-8, 8, (6000, 3)), rng.uniform(-300, 300, (4000, 3))])
    want = simplex3(disk["perm"], pts[:, 0], pts[:, 1], pts[:, 2])
    got = np.array([oracle.simplex(disk["perm"], *p) for p in pts])
    # rounding only (lattice offsets formed in another order; |x| up to 300)
    assert np.max(np.abs(got - want)[:6000]) <= 1e-14 * 64
    assert np.max(np.abs(got - want)) <= 2e-12
    assert 0.3 < np.abs(want).max() <= 1.0 and abs(want.mean()) < 0.05


def _fbm(perm, x, y, z, octaves):
    out, amp, f = np.zeros_like(x), 1.0, 1.0
    for _ in range(octaves):
        out += amp * simplex3(perm, x * f, y * f, z * f)
        amp *= 0.5
        f *= 2.0
    return out


def _inside_points(disk, n, seed, rmin=None):
    rng = np.random.default_rng(seed)
    r_lo = disk["r_isco"] if rmin is None else rmin
    r = rng.uniform(r_lo, disk["r_outer"] - 1e-6, n)
    rr = np.linspace(disk["r_min"], disk["r_outer"], int(disk["n_r"]))
    zm = np.interp(r, rr, disk["z_max"])
    z = rng.uniform(-0.95, 0.95, n) * zm
    ph = rng.uniform(-math.pi, math.pi, n)
    return r, z, ph


def test_fbm_density_modulation(oracle, disk, V, Vcalm):
    """rho = base * max(0, 1 + turbulence * fBm(r cos phi / s, r sin phi / s, z / s)),
    fBm the sum of octaves k of 0.5^k noise(2^k x)."""
    r, z, ph = _inside_points(disk, 2000, 2)
    s = disk["noise_scale"]
    n = _fbm(disk["perm"], r * np.cos(ph) / s, r * np.sin(ph) / s, z / s, int(disk["octaves"]))
    base = np.array([oracle.vol_density(Vcalm, *q) for q in zip(r, z, ph)])
    turb = np.array([oracle.vol_density(V, *q) for q in zip(r, z, ph)])
    assert (base > 0).all()
    want = base * np.maximum(0.0, 1.0 + disk["turbulence"] * n)
    np.testing.assert_allclose(turb, want, rtol=1e-12, atol=1e-14 * base.max())


# ---- density and temperature ---------------------------------------------

def _grid(disk):
    nr, nz = int(disk["n_r"]), int(disk["n_z"])
    rr = disk["r_min"] + (disk["r_outer"] - disk["r_min"]) * np.arange(nr) / (nr - 1)
    return nr, nz, rr


def _taper(disk, r):
    return np.where(r >= disk["r_isco"], 1.0, np.exp(-((disk["r_isco"] - r) / disk["taper_width"]) ** 2))


def test_density_and_temperature_at_nodes(oracle, disk, Vcalm):
    nr, nz, rr = _grid(disk)
    prof = disk["rho_profile"].reshape(nr, nz)
    Tp = disk["T_profile"].reshape(nr, nz)
    checked = 0
    for i in range(3, nr - 3, 37):
        zm = disk["z_max"][i]
        if zm <= 0 or disk["rho_mid"][i] <= 0:
            continue
        for k in (0, 1, 17, 64, 100, nz - 3):
            z = zm * k / (nz - 1)
            rho = oracle.vol_density(Vcalm, rr[i], z, 0.3)
            mid = disk["rho_mid"][i] * disk["rho_scale"] * _taper(disk, rr[i])
            # at a node up to rounding of the node's radius, measured against
            # the midplane value (the profile falls to its 1e-15 floor, and the
            # neighbouring column has its own z_max)
            assert abs(rho - mid * prof[i, k]) <= 1e-9 * mid, (i, k)
            assert abs(oracle.vol_temperature(Vcalm, rr[i], z) - Tp[i, k]) <= 1e-9 * Tp[i, 0], (i, k)
            checked += 1
    assert checked > 40


def test_density_linear_between_radial_nodes_and_taper(oracle, disk, Vcalm):
    """At the midplane the profile is 1, so rho/rho_scale is the radial lerp
    of rho_mid times the taper; temperature is the lerp of the two columns'
    midplane values."""
    nr, nz, rr = _grid(disk)
    Tp = disk["T_profile"].reshape(nr, nz)
    assert np.all(disk["rho_profile"].reshape(nr, nz)[:, 0][disk["rho_mid"] > 0] == 1.0)
    for i in list(range(5, nr - 5, 41)) + [int(np.searchsorted(rr, disk["r_isco"])) - 3]:
        for t in (0.25, 0.5, 0.75):
            r = rr[i] + t * (rr[i + 1] - rr[i])
            want = ((1 - t) * disk["rho_mid"][i] + t * disk["rho_mid"][i + 1]) * disk["rho_scale"] * _taper(disk, r)
            assert oracle.vol_density(Vcalm, r, 0.0, 0.0) == pytest.approx(want, rel=1e-9, abs=1e-300)
            assert oracle.vol_temperature(Vcalm, r, 0.0) == pytest.approx((1 - t) * Tp[i, 0] + t * Tp[i + 1, 0],
                                                                          rel=1e-9)


def test_density_edge_and_zero_regions(oracle, disk, Vcalm):
    """Above z_max: rho(z_max + dz) = rho(0.9999 z_max) exp(-2 (dz/H)^2) up to
    1.5 H, zero beyond, and zero outside (r_horizon, r_outer]; the
    temperature keeps its surface value in the soft edge and is zero beyond."""
    nr, nz, rr = _grid(disk)
    for i in range(20, nr - 5, 53):
        r = rr[i]
        zm, H = disk["z_max"][i], disk["H"][i]
        if zm <= 0:
            continue
        surf = oracle.vol_density(Vcalm, r, zm * 0.9999, 0.0)
        Ts = oracle.vol_temperature(Vcalm, r, zm * 0.9999)
        for f in (0.1, 0.7, 1.4):
            dz = f * H
            # z_max(r) interpolated at the node radius can differ from the
            # table entry by rounding, which moves the 0.9999 z_max lookup
            assert oracle.vol_density(Vcalm, r, zm + dz, 0.0) == pytest.approx(surf * math.exp(-2 * f * f), rel=1e-9)
            assert oracle.vol_density(Vcalm, r, -(zm + dz), 0.0) == pytest.approx(surf * math.exp(-2 * f * f),
                                                                                  rel=1e-9)
            assert oracle.vol_temperature(Vcalm, r, zm + dz) == Ts
        assert oracle.vol_density(Vcalm, r, zm + 1.5 * H * 1.0001, 0.0) == 0.0
        assert oracle.vol_temperature(Vcalm, r, zm + 1.5 * H * 1.0001) == 0.0
    assert oracle.vol_density(Vcalm, disk["r_horizon"], 0.0, 0.0) == 0.0
    assert oracle.vol_density(Vcalm, disk["r_outer"] * 1.0001, 0.0, 0.0) == 0.0
    assert oracle.vol_temperature(Vcalm, disk["r_outer"] * 1.0001, 0.0) == 0.0


# ---- gas velocity --------------------------------------------------------

def _eq_metric(a, r):
    return -(1 - 2 * M / r), -2 * M * a / r, r * r + a * a + 2 * M * a * a / r, r * r / (r * r - 2 * M * r + a * a)


def _bpt_isco(a):
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def test_isco_energy_and_angular_momentum(disk):
    a = disk["a"]
    r = _bpt_isco(a)
    assert disk["r_isco"] == pytest.approx(r, rel=1e-12)
    v = 1 / math.sqrt(r)
    E = (1 - 2 * v * v + a * v ** 3) / math.sqrt(1 - 3 * v * v + 2 * a * v ** 3)
    L = math.sqrt(r) * (1 - 2 * a * v ** 3 + a * a * v ** 4) / math.sqrt(1 - 3 * v * v + 2 * a * v ** 3)
    assert disk["E_isco"] == pytest.approx(E, rel=1e-10)
    assert disk["L_isco"] == pytest.approx(L, rel=1e-10)


def test_gas_velocity_outside_isco_is_keplerian_and_unit(oracle, disk, V):
    a = disk["a"]
    for r in np.linspace(disk["r_isco"], disk["r_outer"], 9):
        ut, ur, up = oracle.vol_velocity(V, r, math.pi / 2)
        gtt, gtp, gpp, _ = _eq_metric(a, r)
        assert ur == 0.0
        assert up / ut == pytest.approx(1.0 / (r ** 1.5 + a), rel=1e-13)
        assert gtt * ut * ut + 2 * gtp * ut * up + gpp * up * up == pytest.approx(-1.0, rel=1e-12)


def test_gas_velocity_inside_isco_conserves_e_and_l(oracle, disk, V):
    a = disk["a"]
    for r in np.linspace(disk["r_horizon"] * 1.05, disk["r_isco"] * 0.999, 9):
        ut, ur, up = oracle.vol_velocity(V, r, math.pi / 2)
        gtt, gtp, gpp, grr = _eq_metric(a, r)
        assert ur < 0.0                                        # infall
        assert -(gtt * ut + gtp * up) == pytest.approx(disk["E_isco"], rel=1e-11)
        assert gtp * ut + gpp * up == pytest.approx(disk["L_isco"], rel=1e-11)
        norm = gtt * ut * ut + 2 * gtp * ut * up + gpp * up * up + grr * ur * ur
        assert norm == pytest.approx(-1.0, rel=1e-9, abs=1e-9)


# ---- opacity lookups -----------------------------------------------------

def _axes(disk):
    n_nu = int(disk["n_nu"])
    lnu = disk["log_nu_min"] + (disk["log_nu_max"] - disk["log_nu_min"]) * np.arange(n_nu) / (n_nu - 1)
    lrho = -18.0 + 12.0 * np.arange(100) / 99
    lT = math.log10(3000.0) + (8.0 - math.log10(3000.0)) * np.arange(100) / 99
    return n_nu, lnu, lrho, lT


def test_kappa_lookups_at_nodes(oracle, disk, V):
    n_nu, lnu, lrho, lT = _axes(disk)
    K = disk["kappa_abs"].reshape(n_nu, 100, 100)
    ES = disk["kappa_es"].reshape(100, 100)
    rng = np.random.default_rng(4)
    for _ in range(300):
        i, j, k = rng.integers(0, n_nu - 1), rng.integers(0, 99), rng.integers(0, 99)
        got = oracle.vol_kappa_abs(V, 10 ** lnu[i], 10 ** lrho[j], 10 ** lT[k])
        assert got == pytest.approx(max(K[i, j, k], 1e-100), rel=1e-9), (i, j, k)
        assert oracle.vol_kappa_es(V, 10 ** lrho[j], 10 ** lT[k]) == pytest.approx(ES[j, k], rel=1e-9)


def test_kappa_lookups_at_cell_midpoints(oracle, disk, V):
    """ln kappa_abs is trilinear in (log nu, log rho, log T): at a cell centre
    it is the mean of the eight corners' ln kappa; kappa_es is bilinear in
    value: the mean of the four corners."""
    n_nu, lnu, lrho, lT = _axes(disk)
    K = np.log(np.maximum(disk["kappa_abs"].reshape(n_nu, 100, 100), 1e-100))
    ES = disk["kappa_es"].reshape(100, 100)
    rng = np.random.default_rng(5)
    for _ in range(300):
        # the last cell of each axis maps to its lower node (the index clamp)
        i, j, k = rng.integers(0, n_nu - 2), rng.integers(0, 98), rng.integers(0, 98)
        nu = 10 ** (0.5 * (lnu[i] + lnu[i + 1]))
        rho = 10 ** (0.5 * (lrho[j] + lrho[j + 1]))
        T = 10 ** (0.5 * (lT[k] + lT[k + 1]))
        want = math.exp(K[i:i + 2, j:j + 2, k:k + 2].mean())
        assert oracle.vol_kappa_abs(V, nu, rho, T) == pytest.approx(want, rel=1e-9), (i, j, k)
        assert oracle.vol_kappa_es(V, rho, T) == pytest.approx(ES[j:j + 2, k:k + 2].mean(), rel=1e-12)


def test_kappa_lookup_clamps(oracle, disk, V):
    """Outside the table the index clamps to the edge cells; at the top node
    the reference's clamp (t <= n - 2) returns the node below it."""
    n_nu, lnu, lrho, lT = _axes(disk)
    K = disk["kappa_abs"].reshape(n_nu, 100, 100)
    j, k = 40, 60
    rho, T = 10 ** lrho[j], 10 ** lT[k]
    assert oracle.vol_kappa_abs(V, 10 ** (lnu[0] - 1.0), rho, T) == pytest.approx(K[0, j, k], rel=1e-9)
    assert oracle.vol_kappa_abs(V, 10 ** lnu[-1], rho, T) == pytest.approx(K[n_nu - 2, j, k], rel=1e-9)
    assert oracle.vol_kappa_abs(V, 10 ** (lnu[-1] + 2.0), rho, T) == pytest.approx(K[n_nu - 2, j, k], rel=1e-9)


# ---- radiative transfer --------------------------------------------------

def test_transfer_step_is_the_formal_solution(oracle):
    """For piecewise-constant sources the backward accumulation is exact:
    J = sum_i S_i e^{-tau_<i} (1 - e^{-dtau_i}), T = e^{-tau}.  Constant S
    gives S (1 - e^{-tau}); an optically thick first sample gives its own S."""
    rng = np.random.default_rng(6)
    dt = rng.exponential(0.05, 200)
    J, T = oracle.rt_steps(np.full(200, 3.7), dt)
    tau = dt.sum()
    assert J == pytest.approx(3.7 * (1 - math.exp(-tau)), rel=1e-13)
    assert T == pytest.approx(math.exp(-tau), rel=1e-12)
    S = rng.lognormal(0, 2, 200)
    before = np.concatenate([[0.0], np.cumsum(dt)[:-1]])
    want = np.sum(S * np.exp(-before) * -np.expm1(-dt))
    J, T = oracle.rt_steps(S, dt)
    assert J == pytest.approx(want, rel=1e-12)
    J, T = oracle.rt_steps([2.5, 9.0], [800.0, 1.0])
    assert J == 2.5 and T == 0.0
    J, T = oracle.rt_steps([5.0, 5.0], [0.0, 0.0])
    assert J == 0.0 and T == 1.0
    # continuing a ray: the state carries over between entries into the volume
    J1, T1 = oracle.rt_steps(S[:70], dt[:70])
    J2, T2 = oracle.rt_steps(S[70:], dt[70:], J=J1, T=T1)
    J0, T0 = oracle.rt_steps(S, dt)
    assert J2 == pytest.approx(J0, rel=1e-14) and T2 == pytest.approx(T0, rel=1e-14)


def test_disk_without_matter_emits_nothing(oracle, disk):
    """rho_scale = 0: every density is zero, so every spectrum is zero and
    no ray is classed as a disk ray, while the geodesics are unchanged."""
    empty = oracle.vol_struct({**disk, "rho_scale": 0.0})
    sc = oracle.scene(width=48, height=48, spin_a=0.9, theta_obs=math.radians(75.0), fov=math.radians(90.0),
                      disk=False)
    ids = np.arange(48 * 48)
    freqs = [10 ** (14 + 4 * k / 31) for k in range(32)]
    rays, spec = oracle.vol_trace_samples(sc, empty, oracle.JITTER_CENTER, 1, 0, ids, freqs=freqs)
    assert np.all(spec == 0.0)
    assert not np.any(rays["cls"] == 2)
    rays_rgb, _ = oracle.vol_trace_samples(sc, empty, oracle.JITTER_CENTER, 1, 0, ids)
    assert np.all(rays_rgb["color"] == 0.0)


# ---- radial structure ----------------------------------------------------

def test_teff_table_against_novikov_thorne_quadrature(disk):
    """T_eff(r) = T_peak (F / F_max)^(1/4) with F the Novikov-Thorne flux
    (its integral by adaptive quadrature with analytic derivatives here),
    and the plunging region at the value just outside the ISCO."""
    from scipy.integrate import quad
    a, ri, ro = disk["a"], disk["r_isco"], disk["r_outer"]

    def E(r):
        w = math.sqrt(M / r ** 3)
        return (1 - 2 * M / r + a * w) / math.sqrt(1 - 3 * M / r + 2 * a * w)

    def L(r):
        w = math.sqrt(M / r ** 3)
        return math.sqrt(M * r) * (1 - 2 * a * w + a * a / r ** 2) / math.sqrt(1 - 3 * M / r + 2 * a * w)

    def Om(r):
        w = math.sqrt(M / r ** 3)
        return w / (1 + a * w)

    def d(f, r, h=1e-4):
        return (f(r - 2 * h) - 8 * f(r - h) + 8 * f(r + h) - f(r + 2 * h)) / (12 * h)

    Ei, Li = E(ri), L(ri)

    def F(r):
        I, _ = quad(lambda x: (E(x) - Ei) * d(L, x) - (L(x) - Li) * d(E, x), ri, r, limit=200)
        return 3 * M / (8 * math.pi * r ** 3) * (-d(Om, r)) * I / (E(r) - Om(r) * L(r))

    rs = np.linspace(ri, ro, 400)[1:]
    Fs = np.array([F(r) for r in rs])
    Fmax = Fs.max()
    nr, _, rr = _grid(disk)
    for i in range(0, nr, 23):
        r = rr[i]
        if r < ri + 0.5 or r >= ro:
            continue
        want = disk["t_peak"] * (F(r) / Fmax) ** 0.25
        assert disk["T_eff"][i] == pytest.approx(want, rel=2e-3), (i, r)
    # Inside the ISCO: the value of the flux table (500 bins from r_isco,
    # trapezoid from 0 at bin 0) linearly interpolated at r_isco + 0.01.
    inside = rr < ri
    dr = (ro - ri) / 499
    r1 = ri + dr
    f1 = (E(r1) - Ei) * d(L, r1) - (L(r1) - Li) * d(E, r1)
    F1 = 3 * M / (8 * math.pi * r1 ** 3) * (-d(Om, r1)) * (0.5 * f1 * dr) / (E(r1) - Om(r1) * L(r1))
    T_plunge = disk["t_peak"] * ((0.01 / dr) * F1 / Fmax) ** 0.25
    assert np.all(disk["T_eff"][inside] == disk["T_eff"][inside][0])
    assert disk["T_eff"][inside][0] == pytest.approx(T_plunge, rel=2e-3)
