"""Output files (no GPU): the FITS spectral cube writer against the FITS
standard's structure (80-character cards, 2880-byte blocks, big-endian IEEE
data, NAXIS3 outermost, bottom row first) and the reference's own check
(tests/test_spectral.cpp:104-139: SIMPLE card first, file of at least 5760
bytes for a 4x3x2 cube); the Radiance RGBE writer decoded by an independent
reader written from the format description (Ward, Graphics Gems II).
"""
import math

import numpy as np
import pytest

BLOCK = 2880


def _cards(raw: bytes):
    cards, off = [], 0
    while True:
        c = raw[off:off + 80].decode("ascii")
        cards.append(c)
        off += 80
        if c.startswith("END "):
            break
    hdr_len = (off + BLOCK - 1) // BLOCK * BLOCK
    assert raw[off:hdr_len] == b" " * (hdr_len - off)
    return cards, hdr_len


def _value(cards, key):
    for c in cards:
        if c[:8].rstrip() == key:
            assert c[8:10] == "= "
            v = c[10:].split(" / ")[0].strip()
            return v.strip("'").rstrip() if v.startswith("'") else v
    raise KeyError(key)


@pytest.fixture(scope="module")
def g():
    import grrt_amd
    return grrt_amd


def test_reference_fits_check(g, tmp_path):
    """The reference's own writer test on the same 4x3x2 cube."""
    data = (np.arange(4 * 3 * 2) * 1.5).reshape(3, 4, 2)
    p = g.make_params(spin=0.998, observer_r=50.0)
    path = str(tmp_path / "t.fits")
    g.write_fits_cube(path, data, [1e14, 1e16], p)
    raw = open(path, "rb").read()
    assert raw[:6] == b"SIMPLE" and "T" in raw[:80].decode()
    assert len(raw) >= 5760 and len(raw) % BLOCK == 0


@pytest.mark.parametrize("bitpix", [-64, -32])
def test_fits_cube_layout(g, tmp_path, bitpix):
    rng = np.random.default_rng(4)
    H, W, N = 9, 17, 5
    cube = rng.lognormal(0, 3, (H, W, N))
    freqs = [10 ** (14 + k) for k in range(N)]            # log-uniform bins
    p = g.make_params(spin=0.9, observer_r=42.0, observer_theta=1.2, fov=0.7, spp=4)
    path = str(tmp_path / "c.fits")
    g.write_fits_cube(path, cube, freqs, p, bitpix=bitpix)
    raw = open(path, "rb").read()
    cards, hdr = _cards(raw)
    assert all(len(c) == 80 for c in cards)
    assert cards[0].startswith("SIMPLE  =                    T")
    assert int(_value(cards, "BITPIX")) == bitpix
    assert [int(_value(cards, k)) for k in ("NAXIS", "NAXIS1", "NAXIS2", "NAXIS3")] == [3, W, H, N]
    assert _value(cards, "BUNIT") == "erg/s/cm2/Hz/sr"
    assert _value(cards, "CTYPE3") == "FREQ-LOG"
    assert float(_value(cards, "CRVAL3")) == pytest.approx(14.0)
    assert float(_value(cards, "CDELT3")) == pytest.approx(1.0)
    assert [float(_value(cards, f"FREQ{k + 1:04d}")) for k in range(N)] == pytest.approx(freqs)
    assert float(_value(cards, "SPIN")) == 0.9 and float(_value(cards, "OBS_R")) == 42.0
    assert int(_value(cards, "SPP")) == 4
    es = 8 if bitpix == -64 else 4
    assert len(raw) == hdr + (H * W * N * es + BLOCK - 1) // BLOCK * BLOCK
    data = np.frombuffer(raw, dtype=">f8" if es == 8 else ">f4", count=H * W * N, offset=hdr)
    planes = data.reshape(N, H, W)[:, ::-1, :]              # FITS row 1 is the bottom
    want = cube if es == 8 else cube.astype(np.float32)
    assert np.array_equal(planes.transpose(1, 2, 0), want)
    assert raw[hdr + H * W * N * es:] == b"\0" * (len(raw) - hdr - H * W * N * es)


def test_fits_frequency_axis_kinds(g, tmp_path):
    cube = np.ones((2, 2, 3))
    for freqs, kind in (([1e14, 2e14, 3e14], "FREQ"), ([1e14, 1e15, 1e16], "FREQ-LOG"),
                        ([1e14, 3e14, 1e16], "FREQ-TAB")):
        path = str(tmp_path / f"{kind}.fits")
        g.write_fits_cube(path, cube, freqs)
        cards, _ = _cards(open(path, "rb").read())
        assert _value(cards, "CTYPE3") == kind
        if kind == "FREQ":
            assert float(_value(cards, "CDELT3")) == pytest.approx(1e14)
            assert not any(c.startswith("FREQ0001") for c in cards)
        else:
            assert float(_value(cards, "FREQ0003")) == pytest.approx(freqs[2])


def _read_rgbe(path):
    """Radiance picture reader: header lines up to a blank line, the
    resolution string, then scanlines either flat RGBE or new-style
    run-length (2, 2, width hi, width lo, then per channel runs: count > 128
    repeats one byte count-128 times, else count literal bytes)."""
    raw = open(path, "rb").read()
    pos = 0
    lines = []
    while True:
        e = raw.index(b"\n", pos)
        line = raw[pos:e].decode()
        pos = e + 1
        if line == "":
            break
        lines.append(line)
    assert lines[0].startswith("#?RADIANCE") and "FORMAT=32-bit_rle_rgbe" in lines
    e = raw.index(b"\n", pos)
    _, h, _, w = raw[pos:e].decode().split()
    h, w = int(h), int(w)
    pos = e + 1
    img = np.zeros((h, w, 4), np.uint8)
    for j in range(h):
        if raw[pos] == 2 and raw[pos + 1] == 2 and 8 <= w < 32768:
            assert (raw[pos + 2] << 8 | raw[pos + 3]) == w
            pos += 4
            for ch in range(4):
                i = 0
                while i < w:
                    n = raw[pos]
                    pos += 1
                    if n > 128:
                        img[j, i:i + n - 128, ch] = raw[pos]
                        pos += 1
                        i += n - 128
                    else:
                        img[j, i:i + n, ch] = np.frombuffer(raw, np.uint8, n, pos)
                        pos += n
                        i += n
        else:
            img[j] = np.frombuffer(raw, np.uint8, 4 * w, pos).reshape(w, 4)
            pos += 4 * w
    assert pos == len(raw)
    ex = img[..., 3].astype(np.int32)
    f = np.where(ex > 0, np.ldexp(1.0, ex - 136), 0.0)     # (m + 0.5) * 2^(e - 128 - 8)
    rgb = np.where(ex[..., None] > 0, (img[..., :3] + 0.5) * f[..., None], 0.0)
    return rgb


@pytest.mark.parametrize("w", [5, 300])
def test_rgbe_round_trip(g, tmp_path, w):
    rng = np.random.default_rng(9)
    h = 7
    fb = np.zeros((h, w, 4), np.float32)
    fb[..., :3] = rng.lognormal(0, 4, (h, w, 3))
    fb[0, 0, :3] = 0.0
    fb[1, 1, :3] = [1e-40, 0.0, 0.0]
    fb[..., 3] = 1.0
    path = str(tmp_path / "a.hdr")
    g.write_hdr(path, fb)
    rgb = _read_rgbe(path)
    m = fb[..., :3].max(-1, keepdims=True).astype(np.float64)
    err = np.abs(rgb - fb[..., :3]) / np.where(m > 1e-30, m, 1.0)
    assert err.max() <= 2.0 / 256
    assert np.all(rgb[0, 0] == 0)
    # normalised: scaled by 0.18 / exp(mean log(1e-6 + L)) over lit pixels
    g.write_hdr(path, fb, normalised=True)
    rgb2 = _read_rgbe(path)
    L = 0.2126 * fb[..., 0] + 0.7152 * fb[..., 1] + 0.0722 * fb[..., 2]
    lit = L > 1e-10
    scale = 0.18 / math.exp(np.mean(np.log(1e-6 + L[lit].astype(np.float64))))
    ref = fb[..., :3].astype(np.float64) * scale
    err2 = np.abs(rgb2 - ref) / np.where(m * scale > 1e-30, m * scale, 1.0)
    assert err2.max() <= 2.0 / 256


# --- the GRRT_BG_TEXTURE reader (host only) -------------------------------

@pytest.mark.parametrize("w", [5, 300])
def test_texture_reader_reads_the_rgbe_writer(g, tmp_path, w):
    """Flat (width < 8) and run-length scanlines: the reader returns what the
    independent decoder above returns, value for value, row 0 first."""
    rng = np.random.default_rng(4)
    h = 6
    fb = np.zeros((h, w, 4), np.float32)
    fb[..., :3] = rng.lognormal(0, 3, (h, w, 3))
    fb[2, :, :3] = 0.0
    fb[..., 3] = 1.0
    path = str(tmp_path / "t.hdr")
    g.write_hdr(path, fb)
    tex = g.read_texture(path)
    ref = _read_rgbe(path).astype(np.float32)
    assert tex.shape == (h, w, 3) and tex.dtype == np.float32
    assert np.array_equal(tex, ref)
    m = fb[..., :3].max(-1, keepdims=True)
    assert (np.abs(tex - fb[..., :3]) / np.where(m > 0, m, 1.0)).max() <= 2.0 / 256
    assert np.all(tex[2] == 0.0)


@pytest.mark.parametrize("little", [True, False])
def test_texture_reader_pfm(g, tmp_path, little):
    from parity_util import write_pfm
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 10, (7, 11, 3)).astype(np.float32)
    path = str(tmp_path / "c.pfm")
    write_pfm(path, img, little=little)
    assert np.array_equal(g.read_texture(path), img)          # bottom-first rows restored
    grey = rng.uniform(0, 1, (4, 9)).astype(np.float32)
    write_pfm(path, grey, little=little)
    assert np.array_equal(g.read_texture(path), np.repeat(grey[..., None], 3, -1))


def test_texture_reader_accepts_crlf_headers(g, tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    path = str(tmp_path / "crlf.pfm")
    with open(path, "wb") as f:
        f.write(b"PF\r\n3 2\r\n-1.0\r\n" + np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())
    assert np.array_equal(g.read_texture(path), img)


def test_texture_reader_clears_negative_and_non_finite(g, tmp_path):
    from parity_util import write_pfm
    img = np.full((2, 3, 3), 0.5, np.float32)
    img[0, 0] = [-1.0, np.nan, np.inf]
    path = str(tmp_path / "bad.pfm")
    write_pfm(path, img)
    tex = g.read_texture(path)
    assert np.all(tex[0, 0] == 0.0) and np.all(tex.reshape(-1, 3)[1:] == 0.5)


def test_texture_reader_errors(g, tmp_path):
    cases = {
        "missing.hdr": None,
        "magic.hdr": b"P6\n2 2\n255\n" + bytes(12),
        "orient.hdr": b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 2 +X 2\n" + bytes(16),
        "short.hdr": b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n" + bytes(10),
        "xyze.hdr": b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n" + bytes(4),
        "short.pfm": b"PF\n2 2\n-1.0\n" + bytes(40),
        "size.pfm": b"PF\n2 x\n-1.0\n" + bytes(48),
    }
    for name, raw in cases.items():
        path = str(tmp_path / name)
        if raw is not None:
            open(path, "wb").write(raw)
        with pytest.raises(RuntimeError, match=path.replace(".", r"\.")):
            g.read_texture(path)
