"""The oracle's particle source (Exposure::sampleParticle) pinned on the CPU:
collimation angles uniform, directions on the pointsFromCollimations geometry
(beamactorcontainer.cpp:44-69) and inside its four field corners, energies
distributed as the spectrum weights (chi-square), weights equal to the exact
piecewise-linear bowtie interpolation of the reference's bowtiefilters.json
nodes; and the tube model's half-value layers against the IEC 61267 standard
radiation qualities."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.optimize import brentq

import opendxmc_amd as dx
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE
from source_checks import (bowtie_nodes, bowtie_weight, check_angles_uniform, check_energy_chi2, corner_rays,
                           direction_of)


def ct_exposure(k=37):
    b = dx.CTSpiralBeam((0, 0, -15), (0, 0, 15), {13: 6.0})
    b.setCollimation(4.0)
    b.setPitch(1.0)
    b.setTubeVoltage(120.0)
    ex = b.exposures()[k:k + 1].copy()
    return b, ex


def test_oracle_source_geometry_and_weights(oracle):
    beam, ex = ct_exposure()
    ex["bowtie_index"] = 0
    ex["weight"] = 0.75
    e, w = beam.spectra()[0]
    nodes = bowtie_nodes(DEFAULT_CT_BOWTIE)
    s = oracle.sample_source(ex, (e, w), nodes, seed=3, n=200_000)
    ax, ay, d, E, wt = s[:, 0], s[:, 1], s[:, 2:5], s[:, 5], s[:, 6]
    half = ex["collimation_half_angles"][0]
    check_angles_uniform(ax, ay, half)
    cos = ex["cosines"][0]
    # the direction of each particle is the collimation geometry at its angles
    np.testing.assert_allclose(d, direction_of(cos, ax, ay), atol=1e-12)
    # inside the field: |d.c0| <= sin(hx), |d.c1| <= sin(hy), forward along c0 x c1
    c0, c1 = cos[:3], cos[3:]
    assert np.all(np.abs(d @ c0) <= math.sin(half[0]) + 1e-12)
    assert np.all(np.abs(d @ c1) <= math.sin(half[1]) + 1e-12)
    assert np.all(d @ np.cross(c0, c1) > 0)
    # the extreme particles approach the four field corners
    corners = corner_rays(cos, half)
    for c in corners:
        assert np.max(d @ c) > math.cos(2e-3)
    check_energy_chi2(E, e, w)
    np.testing.assert_allclose(wt, 0.75 * bowtie_weight(nodes, ax), rtol=1e-12)


def test_oracle_source_dx_field_and_mono(oracle):
    b = dx.DXBeam()
    ex = b.exposures()[:1].copy()
    s = oracle.sample_source(ex, (np.array([60.0]), np.array([1.0])), None, seed=9, n=50_000)
    half = ex["collimation_half_angles"][0]
    check_angles_uniform(s[:, 0], s[:, 1], half)
    np.testing.assert_allclose(s[:, 2:5], direction_of(ex["cosines"][0], s[:, 0], s[:, 1]), atol=1e-12)
    assert np.all(s[:, 5] == 60.0) and np.all(s[:, 6] == ex["weight"][0])


def test_oracle_energy_chi2_detects_a_shifted_spectrum(oracle):
    """The chi-square check is sharp enough to see one bin's weight moved to
    its neighbour on a sample of this size."""
    beam, ex = ct_exposure()
    e, w = beam.spectra()[0]
    s = oracle.sample_source(ex, (e, w), None, seed=4, n=200_000)
    w2 = w.copy()
    j = int(np.argmax(w))
    w2[j + 1] += 0.5 * w2[j]
    w2[j] *= 0.5
    with pytest.raises(AssertionError):
        check_energy_chi2(s[:, 5], e, w2)


# IEC 61267:2005 standard radiation qualities (also IEC 62220-1:2003 Table 2):
# kV, first HVL of the RQR beam (mm Al), added Al of the RQA beam (mm), first
# HVL of the RQA beam (mm Al)
IEC_QUALITIES = [(50, 1.78, 10.0, 4.0), (70, 2.58, 21.0, 7.1), (90, 3.48, 30.0, 9.1), (120, 5.00, 40.0, 11.5)]


@pytest.mark.parametrize("kv,rqr,added,rqa", IEC_QUALITIES)
def test_tube_hvl_iec_61267(kv, rqr, added, rqa):
    """Tube (tungsten, 12 degree anode): the Al filtration that gives the RQR
    beam's HVL is a diagnostic tube's (2.5 to 4.5 mm Al), and adding the RQA
    filter gives the RQA beam's HVL within 10 %.  The RQA HVL is set mainly by
    the thick added filter, so it tests the spectrum shape and the Al
    attenuation together."""
    base = brentq(lambda b: dx.Tube(kv, 12.0, {13: b}).halfValueLayerAl() - rqr, 0.05, 20.0)
    assert 2.5 <= base <= 4.5, base
    hvl = dx.Tube(kv, 12.0, {13: base + added}).halfValueLayerAl()
    assert hvl == pytest.approx(rqa, rel=0.10)
