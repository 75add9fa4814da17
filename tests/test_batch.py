"""Batched transport (dxg_batch_*): instances of uneven size in one persistent
launch give exactly what one dxg_transport_run per instance gives (same RNG
streams keyed by seed and history, so identical event counts and scores up to
the order of FP64 atomic additions), on the first (captured) run and on graph
replays."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import _native as N
from opendxmc_amd import batch_workload as bw


def test_batch_symbols_and_empty_batch():
    b = dx.TransportBatch()
    assert len(b) == 0
    with pytest.raises(N.NativeError) as e:
        b.run()
    assert e.value.code == N.DXG_ERR_STATE


def test_workload_definition():
    inst = bw.instances(8, seed=3, shape_scale=0.25, hist_range=(1e5, 2e5))
    assert len(inst) == 8
    shapes = [ph.dims for ph, _, _ in inst]
    assert shapes[0] == (16, 16, 16) and shapes[1] == (32, 32, 32) and shapes[2] == (24, 40, 20)
    kinds = [type(b).__name__ for _, b, _ in inst]
    assert kinds[0::2] == ["DXBeam"] * 4 and kinds[1::2] == ["CTSequentialBeam"] * 4
    for ph, beam, _ in inst:
        assert 0.05 <= ph.spacing[0] <= 0.2
        assert 0.9e5 <= beam.numberOfParticles() <= 2.1e5


COUNTS = ("histories", "steps", "real_events", "deposits", "uncollided")


@pytest.mark.gpu
def test_batch_equals_one_run_per_instance():
    inst = bw.instances(6, seed=5, shape_scale=0.375, hist_range=(1e5, 6e5))
    single = [bw.make_world(ph) for ph, _, _ in inst]
    batched = [bw.make_world(ph) for ph, _, _ in inst]
    batch = dx.TransportBatch()
    for (ph, beam, s), (world, _) in zip(inst, batched):
        batch.add(world, beam, s)
    assert len(batch) == 6
    for run in range(2):   # the capture, then a graph replay (RNG offsets advance)
        tr = dx.Transport()
        for (ph, beam, s), (world, grid) in zip(inst, single):
            tr.seed = s
            tr(world, beam, None, True)
        ms = batch.run()
        assert ms > 0
        for k, ((_, g1), (_, g2)) in enumerate(zip(single, batched)):
            c1, c2 = g1.counters(), g2.counters()
            for name in COUNTS:
                assert c1[name] == c2[name], (run, k, name, c1[name], c2[name])
            assert g1.histories() == g2.histories()
            e1, e21, n1 = g1.energyScored()
            e2, e22, n2 = g2.energyScored()
            np.testing.assert_array_equal(n1, n2)
            np.testing.assert_allclose(e2, e1, rtol=1e-10, atol=1e-12 * e1.max())
            np.testing.assert_allclose(e22, e21, rtol=1e-10, atol=1e-12 * e21.max())
            assert c2["deposited_energy"] == pytest.approx(c1["deposited_energy"], rel=1e-12)
