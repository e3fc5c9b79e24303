"""Comparison helpers shared by the GPU parity tests."""
import numpy as np


def class_boundary_mask(cls_img: np.ndarray) -> np.ndarray:
    """True where a pixel has a 4-neighbour (or 8) of a different class."""
    c = cls_img
    m = np.zeros(c.shape, bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy == 0 and dx == 0:
                continue
            sh = np.roll(np.roll(c, dy, 0), dx, 1)
            m |= sh != c
    return m


def median_rel_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    sel = (np.abs(b) > floor) & (np.abs(a) > floor)
    if not sel.any():
        return 0.0
    return float(np.median(np.abs(a[sel] - b[sel]) / np.abs(b[sel])))


def photon_orbit_radii(a: float, M: float = 1.0):
    """Prograde and retrograde equatorial photon-orbit radii of Kerr
    (Bardeen, Press & Teukolsky 1972), smaller first."""
    x = min(max(a / M, -1.0), 1.0)
    rp = 2 * M * (1 + np.cos(2.0 / 3.0 * np.arccos(-x)))
    rm = 2 * M * (1 + np.cos(2.0 / 3.0 * np.arccos(x)))
    return min(rp, rm), max(rp, rm)


def critical_score(a, M, p_t, p_th, p_ph, theta, n=None):
    """Minimum over the spherical photon-orbit radii of R(r) / (E^2 r^4), R
    the Kerr radial potential of a null ray with covariant momentum
    (p_t, p_th, p_ph) at polar angle theta: 0 on the critical curve, > 0 for
    rays with no turning point there, < 0 otherwise.  Dense sampling (n
    radii, default 4001), no interpolation: a brute-force reference for the
    library's closed form (the minimum of a quartic in 1/r)."""
    p_t, p_th, p_ph, theta = (np.atleast_1d(np.asarray(v, np.float64)) for v in (p_t, p_th, p_ph, theta))
    r0, r1 = photon_orbit_radii(a, M)
    r = np.linspace(r0, r1, n or 4001)[None, :]
    out = np.empty(len(p_t))
    for i in range(0, len(p_t), 1024):          # bounded memory
        sl = slice(i, i + 1024)
        b = -p_ph[sl] / p_t[sl]
        s = np.maximum(np.abs(np.sin(theta[sl])), 1e-10)
        c = np.cos(theta[sl])
        k = (p_th[sl] / p_t[sl]) ** 2 + c * c * (b * b / (s * s) - a * a) + (b - a) ** 2
        u = r * r + a * a - a * b[:, None]
        out[sl] = ((u * u - (r * r - 2 * M * r + a * a) * k[:, None]) / r ** 4).min(axis=1)
    return out


def write_pfm(path, img, little=True):
    """Portable float map: "PF" (RGB) or "Pf" (grey) for an (h, w, 3) or
    (h, w) array, the scale's sign giving the byte order (negative: little
    endian), rows stored bottom first."""
    img = np.asarray(img, np.float32)
    h, w = img.shape[:2]
    magic = b"PF\n" if img.ndim == 3 else b"Pf\n"
    data = np.ascontiguousarray(img[::-1]).astype("<f4" if little else ">f4")
    with open(path, "wb") as f:
        f.write(magic + f"{w} {h}\n".encode() + (b"-1.0\n" if little else b"1.0\n") + data.tobytes())


def direction_rgb(theta, phi):
    """A smooth colour of the escape direction n = (sin t cos p, sin t sin p,
    cos t): RGB = (1 + (n_z, n_y, n_x)) / 2.  Unchanged by the coordinate
    fold theta -> 2 pi - theta, phi -> phi + pi, so it applies to raw final
    coordinates; R fixes theta, G and B together fix phi."""
    theta, phi = np.asarray(theta, np.float64), np.asarray(phi, np.float64)
    st = np.sin(theta)
    return 0.5 * (1.0 + np.stack([np.cos(theta), st * np.sin(phi), st * np.cos(phi)], -1))


def direction_texture(width, height):
    """direction_rgb at the texel centres of an equirectangular map: row j
    at theta = (j + 1/2) pi / height (row 0 at the north pole), column i at
    phi = -pi + (i + 1/2) 2 pi / width."""
    th = (np.arange(height) + 0.5) * np.pi / height
    ph = -np.pi + (np.arange(width) + 0.5) * 2.0 * np.pi / width
    T, P = np.meshgrid(th, ph, indexing="ij")
    return direction_rgb(T, P).astype(np.float32)
