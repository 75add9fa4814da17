"""dxg_transport_run_static, the transport run that allocates nothing (its
pool, counters and beam tables in the library's fixed per-device scratch),
against dxg_transport_run through the C-ABI: the same seed keys the same
per-history random streams, so event counts are identical and scores equal
up to the order of the FP64 atomic additions.  Two calls in a row share the
scratch, the second with a smaller beam on a smaller grid; a beam whose tables
exceed the scratch is rejected with an error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import _native as N
from opendxmc_amd import phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters
from opendxmc_amd.engine import _Keep, beam_struct
from parity import make_world

pytestmark = pytest.mark.gpu

COUNTS = ("histories", "steps", "real_events", "deposits", "uncollided")


def run(entry: str, ph, beam, seed: int):
    world, grid = make_world(ph, 1)
    k = _Keep()
    b = beam_struct(beam, k, calibrated=False)
    N.check(getattr(N.lib(), entry)(grid._h, C.byref(b), None, 0, seed))
    return world, grid


def ct_case():
    ph = phantoms.ct_body((64, 64, 50), (0.6, 0.6, 0.6), seed=5)
    beam = dx.CTSpiralBeam((0, 0, -15), (0, 0, 15), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setNumberOfParticlesPerExposure(2_000_000 // beam.numberOfExposures())
    return ph, beam


def pencil_case():
    ph = phantoms.water_cube(16, 0.25)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
    beam.setNumberOfParticlesPerExposure(200_000)
    return ph, beam


def test_static_run_equals_transport_run():
    for name, (ph, beam), seed in (("ct spiral 64x64x50", ct_case(), 41), ("pencil 16^3", pencil_case(), 42)):
        print(f"[no-alloc] {name}: {beam.numberOfExposures()} exposures, {beam.numberOfParticles()} histories",
              flush=True)
        _, g_static = run("dxg_transport_run_static", ph, beam, seed)
        _, g_plain = run("dxg_transport_run", ph, beam, seed)
        c1, c2 = g_static.counters(), g_plain.counters()
        for key in COUNTS:
            assert c1[key] == c2[key], (name, key, c1[key], c2[key])
        assert c1["histories"] == beam.numberOfParticles() == g_static.histories()
        e1, e21, n1 = g_static.energyScored()
        e2, e22, n2 = g_plain.energyScored()
        np.testing.assert_array_equal(n1, n2)
        np.testing.assert_allclose(e1, e2, rtol=1e-10, atol=1e-12 * e2.max())
        np.testing.assert_allclose(e21, e22, rtol=1e-10, atol=1e-12 * e22.max())
        assert c1["deposited_energy"] == pytest.approx(c2["deposited_energy"], rel=1e-12)
        assert abs(c1["deposited_energy"] + c1["escaped_energy"] - c1["emitted_energy"]) <= 1e-5 * c1["emitted_energy"]


def test_static_run_rejects_tables_beyond_the_scratch():
    ph, _ = pencil_case()
    world, grid = make_world(ph, 1)
    beam = dx.DXBeam()
    beam.setNumberOfExposures(20_000)   # 20,000 x 88 B of exposure tables > 1 MiB
    beam.setNumberOfParticlesPerExposure(10)
    k = _Keep()
    b = beam_struct(beam, k, calibrated=False)
    rc = N.lib().dxg_transport_run_static(grid._h, C.byref(b), None, 0, 1)
    assert rc == N.DXG_ERR_ARGUMENT
    assert b"scratch" in N.lib().dxg_last_error()
    assert grid.histories() == 0
