"""Statistical parity between the HIP engine and the oracle.

Criteria (stated on the per-voxel energy imparted per history, which is the
dose up to the voxel's fixed rho*V):
  * per voxel, |mean_gpu - mean_oracle| <= 3 combined standard errors for at
    least 99 % of the voxels above 10 % of the maximum (at least one voxel is
    allowed to fail when fewer than 100 voxels qualify, since 3 SE is itself
    a 99.73 % interval);
  * total energy imparted per history agrees within 0.5 % (and within 4
    combined standard errors of the total).
The standard error of a voxel mean over H histories is
sqrt((sum e^2 / H - (sum e / H)^2) / (H - 1)), each deposit counted as an
independent contribution.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import opendxmc_amd as dx
from opendxmc_amd import phantoms


@dataclass
class Verdict:
    n_voxels: int
    n_fail: int
    frac_within: float
    total_gpu: float
    total_oracle: float
    total_rel_diff: float
    total_z: float
    ok_voxels: bool
    ok_total: bool

    @property
    def ok(self) -> bool:
        return self.ok_voxels and self.ok_total

    def __str__(self) -> str:
        return (f"voxels>{'10%'}: {self.n_voxels}, fail {self.n_fail} ({self.frac_within:.4f} within 3SE); "
                f"total per history gpu {self.total_gpu:.6g} oracle {self.total_oracle:.6g} "
                f"rel diff {self.total_rel_diff:+.4%} z {self.total_z:+.2f}")


def voxel_stats(sum_e, sum_e2, h):
    m = sum_e / h
    var = np.maximum(sum_e2 / h - m * m, 0.0) / max(h - 1, 1)
    return m, var


def compare(gpu, orc, h_gpu: int, h_orc: int, threshold: float = 0.10, tol_total: float = 0.005) -> Verdict:
    mg, vg = voxel_stats(gpu[0], gpu[1], h_gpu)
    mo, vo = voxel_stats(orc[0], orc[1], h_orc)
    sel = mo > threshold * mo.max()
    se = np.sqrt(vg + vo)[sel]
    d = np.abs(mg - mo)[sel]
    within = d <= 3.0 * np.where(se > 0, se, np.inf)
    n = int(sel.sum())
    n_fail = int(n - within.sum())
    allowed = max(int(np.floor(0.01 * n)), 1 if n < 100 else 0)
    tg, to = gpu[0].sum() / h_gpu, orc[0].sum() / h_orc
    # SE of the total: treat voxel sums as independent (conservative lower bound on the spread)
    tvar = (np.sum(gpu[1]) / h_gpu - tg * tg) / h_gpu + (np.sum(orc[1]) / h_orc - to * to) / h_orc
    tz = (tg - to) / np.sqrt(max(tvar, 1e-300))
    rel = (tg - to) / to
    return Verdict(n, n_fail, float(within.mean()) if n else 1.0, float(tg), float(to), float(rel), float(tz),
                   n_fail <= allowed, abs(rel) <= tol_total)


def make_world(ph: phantoms.Phantom, level: int = 1):
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=level))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    return world, grid


def beam_inputs(beam):
    ex = beam.exposures()
    spectra = beam.spectra()
    bows = [(b.angle, b.weight) for b in beam.bowties()]
    return ex, spectra, bows


def run_both(ph, beam, level, seed=7, oracle=None, threads=16):
    print(f"[parity] {type(beam).__name__} dims={ph.dims} level={level} histories={beam.numberOfParticles()}", flush=True)
    world, grid = make_world(ph, level)
    ok = dx.Transport(seed=seed)(world, beam, None, True)
    assert ok
    e, e2, _ = grid.energyScored()
    hg = grid.histories()
    counters = grid.counters()
    ex, spectra, bows = beam_inputs(beam)
    se, se2, _, oc = oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, ph.materials, ex, spectra, bows,
                                      level=level, seed=seed + 1000, threads=threads)
    return (e, e2), (se, se2), hg, oc["histories"], counters, oc
