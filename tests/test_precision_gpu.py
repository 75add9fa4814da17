"""FP32 transport against the FP64 oracle at high statistics: the total
energy imparted per history, region sums (the beam axis and shells of the
scatter around it) and every voxel down to 0.1 % of the maximum for a pencil
beam (the scatter, not only the axis, which holds more than half the energy)
and 1 % for a CT spiral.

The GPU keeps photon state and cross-section arithmetic in FP32 and tallies
in FP64; the oracle does everything in FP64.  Each side runs in independent
batches (different seeds); the spread of the batch sums gives an honest
standard error of a total or region sum (deposits of one history are
correlated, so the per-deposit second moments alone would understate it).
Bounds: totals within 0.2 % and 4 standard errors of the difference, region
sums within 4 standard errors (and 1 %); voxels to the parity criterion (99 %
within 3 combined SE)."""
from __future__ import annotations

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters
from parity import batched_pair, check_regions

pytestmark = pytest.mark.gpu

# Shells of the pencil beam's scatter: Chebyshev distance (in voxels) from the
# axis column, which the beam at x = y = 0.05 cm runs through (voxel 32 of 64
# on x and y).
SHELLS = ((0, 0), (1, 1), (2, 3), (4, 7), (8, 15), (16, 32))


def pencil_labels(n=64, axis=32):
    i = np.arange(n)
    r = np.maximum(np.abs(i[None, :, None] - axis), np.abs(i[None, None, :] - axis))   # (1, y, x)
    lab = np.full(r.shape, -1, dtype=np.int64)
    for k, (a, b) in enumerate(SHELLS):
        lab[(r >= a) & (r <= b)] = k
    return np.broadcast_to(lab, (n, n, n)).ravel()


@pytest.mark.parametrize("level", [0, 1, 2])
def test_pencil_fp32_vs_fp64_scatter(oracle, level):
    ph = phantoms.water_cube(64, 0.1)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
    print(f"[precision] pencil 64^3 level {level}", flush=True)
    res = batched_pair(oracle, ph, beam, 1.25e7, 4e6, level, pencil_labels(), len(SHELLS))
    names = [f"shell {a}-{b} voxels off the axis" if b else "axis column" for a, b in SHELLS]
    # voxels above 0.1 % of the maximum: ~2,000, nearly all off the axis
    check_regions(f"pencil level {level}", *res, threshold=1e-3, min_voxels=1500, region_names=names)


def test_ct_spiral_fp32_vs_fp64(oracle):
    ph = phantoms.ct_body((64, 64, 50), (0.6, 0.6, 0.6), seed=5)
    beam = dx.CTSpiralBeam((0, 0, -15), (0, 0, 15), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setStepAngleDeg(15.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    print("[precision] ct spiral 64x64x50", flush=True)
    z = np.repeat(np.arange(ph.dims[2]) * 4 // ph.dims[2], ph.dims[0] * ph.dims[1])   # four slabs along z
    res = batched_pair(oracle, ph, beam, 1.25e7, 2.5e6, 1, z, 4)
    check_regions("ct spiral", *res, threshold=0.01, min_voxels=20000, region_names=[f"z slab {k}" for k in range(4)])
