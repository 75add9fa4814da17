"""The transport kernels' Compton angle sampling (Klein-Nishina by its
tabulated inverse CDF, knCos in device.hpp) against the closed-form
Klein-Nishina CDF, evaluated on the host through dxg_debug_compton_cos (the
kernels' own table and FP32 interpolation; no device needed).

The closed form is pinned first against a direct numerical integration of the
textbook cross-section dsigma/dOmega ~ P^2 (P + 1/P - sin^2 theta), P = E'/E
(Klein & Nishina 1929), so a dropped term or a wrong sign in either fails.
The table's bound (5e-5 in probability) is four times tighter than what a
table row shifted by a quarter keV gives (1.9e-4 at 60 keV).

Correction levels 1 and 2 sample Klein-Nishina x S(x)/Z from one table per
material (knsLookup, dxg_debug_compton_cos_material), checked against the CDF
integrated with the oracle's own S(x) (the oracle is the checker here)."""
from __future__ import annotations

import numpy as np
import pytest

from opendxmc_amd.engine import compton_cos

ME_C2 = 510.99895   # keV


def kn_cdf(c, e_kev):
    """P(cos theta <= c) under Klein-Nishina at e_kev, in closed form: with
    x = E/E' = 1 + k (1 - c) the density per unit x is
    x^-3 + (2/k + 1/k^2) x^-2 + (1 - 2/k - 2/k^2) x^-1 + 1/k^2."""
    k = e_kev / ME_C2

    def anti(x):
        return -0.5 / x**2 - (2 / k + 1 / k**2) / x + (1 - 2 / k - 2 / k**2) * np.log(x) + x / k**2

    x_back = 1 + 2 * k
    return (anti(x_back) - anti(1 + k * (1 - np.asarray(c, dtype=np.float64)))) / (anti(x_back) - anti(1.0))


def kn_cdf_numeric(c, e_kev, n=200001):
    k = e_kev / ME_C2
    cc = np.linspace(-1.0, 1.0, n)
    P = 1.0 / (1.0 + k * (1.0 - cc))
    pdf = P * P * (P + 1.0 / P - (1.0 - cc * cc))
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(cc))])
    return np.interp(c, cc, cum / cum[-1])


ENERGIES = [1.0, 1.5, 10.0, 30.3, 60.0, 117.7, 150.0, 199.9, 200.0]


@pytest.mark.parametrize("e", [1.0, 30.0, 120.0, 200.0])
def test_closed_form_matches_the_textbook_cross_section(e):
    c = np.linspace(-1.0, 1.0, 41)
    np.testing.assert_allclose(kn_cdf(c, e), kn_cdf_numeric(c, e), atol=1e-7)


def test_thomson_limit():
    # k -> 0: density (1 + c^2), CDF (c + c^3/3 + 4/3) / (8/3), so F(0) = 1/2
    # and F(-1/2) = 19/64; the closed form cancels too much at tiny k, the
    # numerical integral does not
    assert abs(kn_cdf_numeric(0.0, 1e-3) - 0.5) < 1e-5
    assert abs(kn_cdf_numeric(-0.5, 1e-3) - 19 / 64) < 1e-5
    # at 1 keV the closed form is already near it, and backscatter is favoured
    # less as the energy grows
    assert abs(kn_cdf(0.0, 1.0) - 0.5) < 3e-3
    assert kn_cdf(0.0, 200.0) < kn_cdf(0.0, 50.0) < kn_cdf(0.0, 1.0)


@pytest.mark.parametrize("e", ENERGIES)
def test_table_inverts_the_klein_nishina_cdf(e):
    u = np.linspace(0.0, 1.0, 40001, endpoint=False).astype(np.float32)
    c = compton_cos(e, u).astype(np.float64)
    assert c[0] == -1.0
    assert np.all(c >= -1.0) and np.all(c <= 1.0)
    assert np.all(np.diff(c) >= 0.0)
    err = np.abs(kn_cdf(c, e) - u.astype(np.float64))
    assert err.max() < 5e-5, f"{e} keV: max |F(cos) - u| = {err.max():.2e}"


def test_table_over_the_energy_range():
    # energies between the table's rows (4 keV apart) and across the whole range
    u = np.linspace(0.0, 1.0, 4001, endpoint=False).astype(np.float32)
    worst = 0.0
    for e in np.append(np.arange(1.0, 200.0, 0.37), 200.0):
        c = compton_cos(e, u).astype(np.float64)
        worst = max(worst, np.abs(kn_cdf(c, e) - u.astype(np.float64)).max())
    assert worst < 5e-5, worst


def test_a_shifted_row_would_fail():
    # the bound above is tight enough to see a quarter-keV error in the row
    u = np.linspace(0.0, 1.0, 40001, endpoint=False).astype(np.float32)
    c = compton_cos(60.25, u).astype(np.float64)
    assert np.abs(kn_cdf(c, 60.0) - u).max() > 1e-4


def test_mean_energy_transfer():
    # <1 - E'/E> over the table's angles against the same integral of the
    # closed-form density (a check of the whole distribution, not one point)
    for e in (20.0, 80.0, 150.0):
        k = e / ME_C2
        u = (np.arange(200000, dtype=np.float64) + 0.5) / 200000
        c = compton_cos(e, u.astype(np.float32)).astype(np.float64)
        got = np.mean(1.0 - 1.0 / (1.0 + k * (1.0 - c)))
        cc = np.linspace(-1.0, 1.0, 400001)
        P = 1.0 / (1.0 + k * (1.0 - cc))
        pdf = P * P * (P + 1.0 / P - (1.0 - cc * cc))
        want = np.trapezoid((1.0 - P) * pdf, cc) / np.trapezoid(pdf, cc)
        assert abs(got / want - 1.0) < 2e-4


def test_rejects_bad_arguments():
    with pytest.raises(Exception):
        compton_cos(0.5, [0.1])
    with pytest.raises(Exception):
        compton_cos(60.0, [1.0])


# ---- correction levels 1 and 2: Klein-Nishina x S(x)/Z per material

def _kns_cdf(oracle, comp, e_kev, n=400001):
    """CDF of cos(theta) under KN(cos) S(x)/Z, x = E sin(theta/2) / hc, with
    S(x)/Z from the oracle's own reading of the physics list, integrated by the
    trapezoid rule on a fine grid."""
    c = np.linspace(-1.0, 1.0, n)
    k = e_kev / ME_C2
    P = 1.0 / (1.0 + k * (1.0 - c))
    kn = P * P * (P + 1.0 / P - (1.0 - c * c))
    x = e_kev / 12.398419843320026 * np.sqrt(np.maximum(0.0, 0.5 * (1.0 - c)))
    s, _, z = oracle.scatter_functions(comp, x)
    pdf = kn * np.minimum(1.0, s / z)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(c))])
    return c, cum / cum[-1]


WATER = {1: 0.111894, 8: 0.888106}
MATERIALS = {"water": WATER, "bone-like CaO": {20: 0.4, 8: 0.6}, "iodine": {53: 1.0}}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_material_table_inverts_kn_times_scatter_function(oracle, name):
    """The kernels' per-material table (rows 12 per octave, nodes denser
    towards forward scatter) against the CDF integrated from the oracle's
    S(x): within 1.5e-4 over 1-200 keV, row energies and between them."""
    comp = MATERIALS[name]
    u = np.linspace(0.0, 1.0, 8001, endpoint=False).astype(np.float32)
    worst = 0.0
    for e in np.append(np.exp(np.linspace(0.0, np.log(200.0), 60)), [11.0, 19.0, 20.0, 61.7]):
        c = compton_cos(float(e), u, comp).astype(np.float64)
        assert c[0] == -1.0 and np.all(np.diff(c) >= 0.0) and c.max() <= 1.0
        cc, cum = _kns_cdf(oracle, comp, float(e))
        worst = max(worst, np.abs(np.interp(c, cc, cum) - u).max())
    assert worst < 1.5e-4, worst


def test_scatter_function_shapes_the_material_table(oracle):
    """S(x) suppresses forward scattering, more for low energies and heavy
    atoms: the table differs from Klein-Nishina alone, and the water table
    would fail the iodine CDF (a material mix-up is seen)."""
    u = np.linspace(0.0, 1.0, 4001, endpoint=False).astype(np.float32)
    e = 30.0
    kn = compton_cos(e, u).astype(np.float64)
    w = compton_cos(e, u, WATER).astype(np.float64)
    i = compton_cos(e, u, {53: 1.0}).astype(np.float64)
    assert np.mean(w) < np.mean(kn) and np.mean(i) < np.mean(w)
    cc, cum = _kns_cdf(oracle, {53: 1.0}, e)
    assert np.abs(np.interp(w, cc, cum) - u).max() > 1e-2
