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


BATCHES = 8   # independent batches per side in batched_pair


def batched(run_batch, n_batches, labels, n_regions):
    """Sum of (e, e2) and histories over batches, and per batch the total and
    the region sums (voxels labelled 0..n_regions-1), per history."""
    e = e2 = None
    h = 0
    sums = []
    for b in range(n_batches):
        be, be2, bh = run_batch(b)
        sel = labels >= 0
        regions = np.bincount(labels[sel], weights=be[sel], minlength=n_regions)
        sums.append(np.concatenate(([be.sum()], regions)) / bh)
        e = be if e is None else e + be
        e2 = be2 if e2 is None else e2 + be2
        h += bh
    return (e, e2), h, np.array(sums)


def batched_pair(oracle, ph, beam, gpu_per_batch, oracle_per_batch, level, labels, n_regions, threads=16):
    """GPU and oracle runs of one beam in BATCHES independent batches each (other
    seeds); returns the summed scores, the histories and the per-batch total
    and region sums per history of both sides."""
    world, grid = make_world(ph, level)
    n_ex = beam.numberOfExposures()
    assert labels.shape == (int(np.prod(ph.dims)),)

    def gpu_batch(b):
        grid.clearScores()
        beam.setNumberOfParticlesPerExposure(int(gpu_per_batch // n_ex))
        assert dx.Transport(seed=501 + b)(world, beam, None, False)
        e, e2, _ = grid.energyScored()
        return e, e2, grid.histories()

    ex, spectra, bows = beam_inputs(beam)
    ex = ex.copy()
    ex["n_particles"] = int(oracle_per_batch // len(ex))

    def oracle_batch(b):
        se, se2, _, oc = oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, ph.materials, ex, spectra,
                                          bows, level=level, seed=9001 + b, threads=threads)
        return se, se2, oc["histories"]

    g, hg, sg = batched(gpu_batch, BATCHES, labels, n_regions)
    o, ho, so = batched(oracle_batch, BATCHES, labels, n_regions)
    return g, o, hg, ho, sg, so


def check_regions(tag, g, o, hg, ho, sg, so, threshold, min_voxels, region_names=(), tol_total=0.002, tol_region=0.01):
    """Voxels above threshold x max to the parity criterion; the total and each
    region sum within tol and 4 standard errors of the difference, the
    standard errors from the spread of the batch sums.  g and o may be
    coarsened (coarsen) where single voxels hold too few events."""
    v = compare(g, o, hg, ho, threshold=threshold)
    diff = sg.mean(axis=0) - so.mean(axis=0)
    se = np.sqrt(sg.var(axis=0, ddof=1) / len(sg) + so.var(axis=0, ddof=1) / len(so))
    rel, z = diff / so.mean(axis=0), diff / se
    print(f"[regions] {tag}: voxels >{threshold:.1%} of max: {v.n_voxels}, outside 3 SE {v.n_fail}; total per "
          f"history gpu {sg[:, 0].mean():.6g} oracle {so[:, 0].mean():.6g} keV, rel diff {rel[0]:+.4%} "
          f"+- {se[0] / so[:, 0].mean():.4%} (z {z[0]:+.2f}); histories {hg:.3g} / {ho:.3g}", flush=True)
    for k, name in enumerate(region_names, start=1):
        print(f"[regions] {tag}: {name}: {so[:, k].mean() / so[:, 0].mean():.2%} of the energy, rel diff "
              f"{rel[k]:+.4%} +- {se[k] / so[:, k].mean():.4%} (z {z[k]:+.2f})", flush=True)
    assert v.n_voxels >= min_voxels
    assert v.ok_voxels, str(v)
    assert abs(rel[0]) < tol_total and abs(z[0]) < 4.0, (rel[0], z[0])
    for k, name in enumerate(region_names, start=1):
        assert abs(rel[k]) < tol_region and abs(z[k]) < 4.0, (name, rel[k], z[k])


def coarsen(scores, dims, f):
    """(sum e, sum e^2) summed over blocks of f x f x f voxels (dims x, y, z,
    each a multiple of f); deposits stay independent, so the second moments add."""
    nx, ny, nz = dims
    return tuple(a.reshape(nz // f, f, ny // f, f, nx // f, f).sum(axis=(1, 3, 5)).ravel() for a in scores)
