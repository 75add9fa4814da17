"""The transport kernels' own particle source (the refill path's
sampleSource, replayed by dxg_debug_source) against what the reference fixes:
exposure of each history, collimation angles uniform, directions on the
pointsFromCollimations geometry (beamactorcontainer.cpp:44-69), energies
distributed as the spectrum weights, weights equal to the bowtie
interpolation of the bowtiefilters.json nodes (bowtiefilterreader.cpp:51-97)
times the exposure weight."""
from __future__ import annotations

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters
from opendxmc_amd.engine import sample_source
from source_checks import bowtie_nodes, bowtie_weight, check_angles_uniform, check_energy_chi2, direction_of

pytestmark = pytest.mark.gpu


def small_world():
    ph = phantoms.water_cube(8, 0.5)
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    return world


def check_beam(world, beam, n, nodes_by_bowtie):
    print(f"[source] {type(beam).__name__} n={n}", flush=True)
    ex = beam.exposures()
    s = sample_source(world, beam, n, seed=17)
    ax, ay, d, E, w, ei = s[:, 0], s[:, 1], s[:, 2:5], s[:, 5], s[:, 6], s[:, 7].astype(np.int64)
    # the exposure of beam history h: consecutive blocks of n_particles
    starts = np.concatenate([[0], np.cumsum(ex["n_particles"].astype(np.int64))])
    h = np.arange(n)
    np.testing.assert_array_equal(ei, np.searchsorted(starts, h, side="right") - 1)
    halves = ex["collimation_half_angles"][ei]
    for half in np.unique(halves, axis=0):   # e.g. the two tubes' fields of a dual-source beam
        m = np.all(halves == half, axis=1)
        check_angles_uniform(ax[m].astype(np.float64), ay[m].astype(np.float64), half)
    cos = ex["cosines"][ei]
    ref = np.empty_like(d, dtype=np.float64)
    for k in np.unique(ei):
        m = ei == k
        ref[m] = direction_of(cos[m][0], ax[m].astype(np.float64), ay[m].astype(np.float64))
    np.testing.assert_allclose(d, ref, atol=2e-5)
    spectra = beam.spectra()
    for si, (e, wt) in enumerate(spectra):
        m = ex["spectrum_index"][ei] == si
        check_energy_chi2(E[m], e, wt)
    bw = np.ones(n)
    for bi, nodes in nodes_by_bowtie.items():
        m = ex["bowtie_index"][ei] == bi
        bw[m] = bowtie_weight(nodes, ax[m].astype(np.float64))
    np.testing.assert_allclose(w, ex["weight"][ei] * bw, rtol=2e-5)


def test_source_ct_spiral_bowtie():
    world = small_world()
    beam = dx.CTSpiralBeam((0.0, 0.0, -15.0), (0.0, 0.0, 15.0), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setTubeVoltage(120.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setAECFilter(dx.CTAECFilter((0, 0, -15), (0, 0, 15), (1.0, 2.0, 0.5)))
    beam.setNumberOfParticlesPerExposure(5000)
    check_beam(world, beam, 2_000_000, {0: bowtie_nodes(DEFAULT_CT_BOWTIE)})


def test_source_dual_source_two_spectra():
    world = small_world()
    beam = dx.CTSpiralDualEnergyBeam((0.0, 0.0, -8.0), (0.0, 0.0, 8.0), {13: 6.0})
    beam.setTubeAVoltage(80.0)
    beam.setTubeBVoltage(140.0)
    beam.setCollimation(4.0)
    beam.setPitch(3.2)
    bow = read_filters()[DEFAULT_CT_BOWTIE]
    beam.setBowtieFilterA(bow)
    beam.setBowtieFilterB(bow)
    beam.setNumberOfParticlesPerExposure(20000)
    check_beam(world, beam, 1_500_000, {0: bowtie_nodes(DEFAULT_CT_BOWTIE), 1: bowtie_nodes(DEFAULT_CT_BOWTIE)})


class _UnevenBeam:
    """A DX beam whose exposures hold different particle counts (the
    exposure lookup's search path instead of the h / ppe path)."""

    def __init__(self):
        self.b = dx.DXBeam()
        self.b.setNumberOfExposures(7)
        self.b.setNumberOfParticlesPerExposure(1000)

    def exposures(self):
        ex = self.b.exposures().copy()
        ex["n_particles"] = [1000, 1, 250_000, 3, 40_000, 100_000, 77]
        return ex

    def spectra(self):
        return self.b.spectra()

    def bowties(self):
        return []

    def calibrationFactor(self):
        return 0.0


def test_source_uneven_exposures():
    world = small_world()
    beam = _UnevenBeam()
    n = int(beam.exposures()["n_particles"].sum())
    check_beam(world, beam, n, {})
