"""The physics list generator against closed forms and published known answers.

Pins (each chosen so a plausible slip -- a dropped factor, a wrong power, a
swapped quadrature weight -- fails one of them):
  * hydrogen's form factor is exactly 1/(1 + (q a0 / 2)^2)^2 (Slater 1s, Z=1);
  * F(0) = Z, S(0) = 0, S(large x) -> Z for every element;
  * Slater's rules reproduce the textbook effective charges (C 2p 3.25,
    O 2p 4.55, Fe 3d 6.25, Zn 4s 4.35);
  * Stobbe's hydrogen photo-ionisation cross section at threshold is
    6.30e-18 cm^2 and falls as E^-3.5 far above it;
  * the angular quadrature integrates Klein-Nishina to its closed-form total
    and Thomson to 8 pi r_e^2 / 3;
  * known answers: NIST (Hubbell & Seltzer) water mu/rho 0.2059 (60 keV),
    0.1837 (80), 0.1707 (100), 0.1505 (150) cm2/g within 1 %, 0.3756 (30)
    and 0.2683 (40) within 6 %; aluminium 0.1704 cm2/g at 100 keV within 1.5 %.
    The photoelectric model is screened-hydrogenic, so below ~40 keV it runs
    5-15 % low against EPDL-derived data (the reference's DXMClib data is not
    available; that gap is recorded as unpinned).
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "opendxmc_amd" / "physics"))
import gen_physics_list as g  # noqa: E402

NA = 6.02214076e23


def mu_rho(comp: dict[int, float], energies) -> np.ndarray:
    e = np.asarray(energies, dtype=float)
    tot = np.zeros_like(e)
    s = sum(comp.values())
    for z, w in comp.items():
        inc, coh = g.scatter_cross_sections(z, e)
        pe = g.photo_element(z, e)
        tot += w / s * NA / g.ATOMIC_MASS[z - 1] * 1e-24 * (pe + inc + coh)
    return tot


def test_hydrogen_form_factor_closed_form():
    x = np.linspace(0.0, 5.0, 101)
    F, S = g.form_factors(1, x)
    q = 4 * math.pi * x * g.A0_ANG
    exact = 1.0 / (1.0 + (q / 2.0) ** 2) ** 2
    np.testing.assert_allclose(F, exact, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(S, 1.0 - exact ** 2, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("z", [1, 6, 8, 20, 53, 74, 92])
def test_form_factor_limits(z):
    F, S = g.form_factors(z, np.array([0.0, 200.0]))
    assert F[0] == pytest.approx(z, rel=1e-12)
    assert S[0] == pytest.approx(0.0, abs=1e-9)
    assert F[1] < 1e-2 * z
    assert S[1] == pytest.approx(z, rel=1e-3)


@pytest.mark.parametrize("z,n,l,expected", [(6, 2, 1, 3.25), (8, 2, 1, 4.55), (26, 3, 2, 6.25), (30, 4, 0, 4.35)])
def test_slater_rules_textbook(z, n, l, expected):
    assert g.slater_zeff(z, g.configuration(z), n, l) == pytest.approx(expected, abs=1e-9)


def test_configuration_counts_electrons():
    for z in range(1, g.ZMAX + 1):
        assert sum(occ for _, _, occ in g.configuration(z)) == z


def test_stobbe_hydrogen_threshold_and_tail():
    b = g.K_EDGE_KEV[0]
    thr = g.photo_element(1, np.array([b * (1 + 1e-9)]))[0] * 1e-24
    assert thr == pytest.approx(6.30e-18, rel=3e-3)
    # the non-relativistic formula's asymptote: corrections are O(sqrt(I/E))
    e = np.array([1e4, 2e4])
    s = g.photo_element(1, e)
    slope = math.log(s[1] / s[0]) / math.log(2.0)
    assert slope == pytest.approx(-3.5, abs=0.005)


@pytest.mark.parametrize("e_kev", [10.0, 60.0, 150.0, 500.0])
def test_quadrature_klein_nishina_total(e_kev):
    v, w = g._angular_nodes()
    mu = 1 - 2 * v * v
    k = e_kev / g.ME_C2_KEV
    num = np.sum(g.kn_dcs(k, mu) * 2 * math.pi * 4 * v * w)
    assert num == pytest.approx(g.kn_total(k), rel=1e-9)


def test_quadrature_thomson_total():
    v, w = g._angular_nodes()
    mu = 1 - 2 * v * v
    num = np.sum(0.5 * g.R_E_CM ** 2 * (1 + mu * mu) * 2 * math.pi * 4 * v * w)
    assert num == pytest.approx(8 * math.pi / 3 * g.R_E_CM ** 2, rel=1e-12)


def test_incoherent_approaches_klein_nishina():
    # at 200 keV binding is negligible for hydrogen: sigma_inc -> Z sigma_KN
    inc, _ = g.scatter_cross_sections(1, np.array([200.0]))
    assert inc[0] * 1e-24 == pytest.approx(g.kn_total(200.0 / g.ME_C2_KEV), rel=2e-3)


WATER = {1: 0.111894, 8: 0.888106}


@pytest.mark.parametrize("e,nist,tol", [(60, 0.2059, 0.01), (80, 0.1837, 0.01), (100, 0.1707, 0.01),
                                        (150, 0.1505, 0.01), (40, 0.2683, 0.06), (30, 0.3756, 0.06)])
def test_water_known_answers(e, nist, tol):
    assert mu_rho(WATER, [e])[0] == pytest.approx(nist, rel=tol)


def test_aluminium_known_answer():
    assert mu_rho({13: 1.0}, [100.0])[0] == pytest.approx(0.1704, rel=0.015)


def test_k_edge_jump():
    # calcium: the empirical K jump ratio 125/Z + 3.5 = 9.75 is met within 5 %;
    # for high Z the model's L shells are too weak near threshold (iodine ~10
    # against ~6 measured), a recorded limitation of the hydrogenic L shells
    eb = g.K_EDGE_KEV[19]
    lo, hi = g.photo_element(20, np.array([eb * (1 - 1e-6), eb * (1 + 1e-6)]))
    assert hi / lo == pytest.approx(125 / 20 + 3.5, rel=0.05)
    eb = g.K_EDGE_KEV[52]
    lo, hi = g.photo_element(53, np.array([eb * (1 - 1e-6), eb * (1 + 1e-6)]))
    assert 4.0 < hi / lo < 12.0


def test_table_edges_are_sharp():
    # the doubled edge node carries the value below, then the value above
    e, v = g.photo_table(53, g.energy_grid())
    k = g.photo_k_table(53, e)
    i = np.where(np.abs(e - g.K_EDGE_KEV[52]) < 1e-12)[0]
    assert len(i) == 2 and v[i[1]] / v[i[0]] > 4.0
    assert k[i[0]] == 0.0 and k[i[1]] > 0.7 * v[i[1]]
    assert np.all(k[e < g.K_EDGE_KEV[52]] == 0.0)


def test_k_fluorescence_data():
    # yields of the semi-empirical fit at textbook points (Ca ~0.16, I ~0.88)
    assert g.fluorescence_yield_k(20) == pytest.approx(0.163, abs=0.01)
    assert g.fluorescence_yield_k(53) == pytest.approx(0.88, abs=0.01)
    om, ka1, ka2, kb, p1, p2, p3 = g.k_lines(74)
    assert ka1 == pytest.approx(59.318, abs=1e-3) and ka2 == pytest.approx(57.981, abs=1e-3)
    assert p1 + p2 + p3 == pytest.approx(1.0) and p1 == pytest.approx(2 * p2)
    assert g.k_lines(8)[0] == 0.0   # oxygen's K line is below the 1 keV cutoff


def test_written_file_roundtrip(tmp_path):
    p = tmp_path / "pl.bin"
    g.write(p, zmax=8)
    data = p.read_bytes()
    assert data[:8] == b"DXGPHYS2"
    zmax, ne, nx = np.frombuffer(data[8:20], dtype="<u4")
    assert (zmax, ne, nx) == (8, g.N_ENERGY, g.N_X)
