"""The reference's own opacity known-answer checks (tests/test_opacity.cpp),
kept as data in tests/golden/reference_opacity_kats.json and run against
libgrrt.so's host-side Saha solver and point opacities (no GPU)."""
import json
import math
import pathlib

import pytest

KATS = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_opacity_kats.json").read_text())


def _value(spec, rho):
    """A number, or {"n_e_full": f}: f times the fully ionised electron
    density rho (1 + X) / (2 m_p), X = 0.70."""
    if isinstance(spec, (int, float)):
        return float(spec)
    return spec["n_e_full"] * rho * (1.0 + 0.70) / (2.0 * KATS["m_p"])


@pytest.fixture(scope="module")
def g():
    import grrt_amd
    return grrt_amd


@pytest.mark.parametrize("kat", KATS["saha"], ids=lambda k: f"{k['cite']}:{k['field']}")
def test_saha_kat(g, kat):
    s = g.saha(kat["rho"], kat["T"])
    got = s["n_Hminus"] / s["n_e"] if kat["field"] == "n_Hminus/n_e" else s[kat["field"]]
    if "expect" in kat:
        assert got == pytest.approx(_value(kat["expect"], kat["rho"]), rel=kat["rel_tol"])
    if "below" in kat:
        assert got < _value(kat["below"], kat["rho"])
    if "at_least" in kat:
        assert got >= kat["at_least"]
    if "above" in kat:
        assert got > kat["above"]


@pytest.mark.parametrize("kat", KATS["opacity"], ids=lambda k: f"{k['cite']}:{k['field']}")
def test_opacity_kat(g, kat):
    o = g.opacity(kat["nu"], kat["rho"], kat["T"])
    got = o[kat["field"]]
    if "expect" in kat:
        assert got == pytest.approx(kat["expect"], rel=kat["rel_tol"])
    if kat.get("finite_positive"):
        assert math.isfinite(got) and got > 0.0
    if kat.get("finite_non_negative"):
        assert math.isfinite(got) and got >= 0.0
    if "greater_than_field" in kat:
        assert got > o[kat["greater_than_field"]]
