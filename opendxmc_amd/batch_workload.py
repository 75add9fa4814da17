"""Batched uneven workload: many independent instances (phantom + beam) run
through one TransportBatch launch.

    python -m opendxmc_amd.batch_workload --instances 64 --runs 3

Instance i: a seeded ellipsoid phantom (phantoms.ellipsoid) whose shape
cycles through 64^3, 128^3, 96x160x80 and 256x256x120 voxels at a spacing
drawn from U[0.05, 0.2] cm; even i get a DX projection (70-120 kVp, 2 mm Al +
0.1 mm Cu, square field 20-40 cm), odd i a CT sequential beam (reference
defaults: 9 mm Al, 3.84 cm collimation); histories log-uniform in [1e6, 2e7].
Prints histories/s of the batched launch and of the same instances run one
`Transport` call at a time.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import engine as dx
from . import phantoms

SHAPES = ((64, 64, 64), (128, 128, 128), (96, 160, 80), (256, 256, 120))


def instances(n: int = 64, seed: int = 1, shape_scale: float = 1.0, hist_range=(1e6, 2e7)):
    """[(phantom, beam, seed)] for n instances."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = tuple(max(8, int(round(s * shape_scale))) for s in SHAPES[i % len(SHAPES)])
        spacing = float(rng.uniform(0.05, 0.2))
        ph = phantoms.ellipsoid(shape, (spacing,) * 3, seed=seed * 1000 + i)
        h = float(np.exp(rng.uniform(math.log(hist_range[0]), math.log(hist_range[1]))))
        if i % 2 == 0:
            beam = dx.DXBeam({13: 2.0, 29: 0.1})
            beam.tube.voltage = float(rng.uniform(70.0, 120.0))
            f = float(rng.uniform(20.0, 40.0))
            beam.setCollimation((f, f))
            beam.setNumberOfParticlesPerExposure(int(h))
        else:
            beam = dx.CTSequentialBeam((0.0, 0.0, 0.0))
            beam.setNumberOfSlices(2)
            beam.setSliceSpacing(3.84)
            beam.setNumberOfParticlesPerExposure(max(1, int(h / beam.numberOfExposures())))
        out.append((ph, beam, 100 + i))
    return out


def make_world(ph, level: int = 1):
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=level))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    return world, grid


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="batched uneven workload")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-single", action="store_true", help="skip the one-call-per-instance comparison")
    args = ap.parse_args(argv)

    inst = instances(args.instances, args.seed)
    worlds = [make_world(ph) for ph, _, _ in inst]
    voxels = sum(ph.size for ph, _, _ in inst)
    batch = dx.TransportBatch()
    for (ph, beam, s), (world, _) in zip(inst, worlds):
        batch.add(world, beam, s)
    hist = sum(beam.numberOfParticles() for _, beam, _ in inst)
    print(f"{len(inst)} instances, {voxels / 1e6:.1f} M voxels, {hist:.3e} histories per run", flush=True)
    res = {"instances": len(inst), "voxels": voxels, "histories_per_run": hist, "batched": []}
    for r in range(args.runs):
        t0 = time.perf_counter()
        ms = batch.run()
        wall = time.perf_counter() - t0
        res["batched"].append({"kernel_ms": ms, "wall_s": wall})
        print(f"batched run {r}: kernel {ms:.1f} ms ({hist / (ms * 1e-3):.3e} hist/s), wall {wall:.2f} s", flush=True)
    if not args.no_single:
        tr = dx.Transport()
        t0 = time.perf_counter()
        kms = 0.0
        for (ph, beam, s), (world, grid) in zip(inst, worlds):
            tr.seed = s
            tr(world, beam, None, True)
            kms += grid.counters()["kernel_ms"]
        wall = time.perf_counter() - t0
        res["single"] = {"kernel_ms": kms, "wall_s": wall}
        print(f"one call per instance: kernel {kms:.1f} ms ({hist / (kms * 1e-3):.3e} hist/s), wall {wall:.2f} s",
              flush=True)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
