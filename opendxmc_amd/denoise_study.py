"""Denoiser workflow on one MI355X: engine-generated training pairs, training,
and the time-to-error comparison of Monte Carlo + denoiser against plain
Monte Carlo.

  python -m opendxmc_amd.denoise_study --out result.json [options]

1. Training data.  Seeded CT body phantoms (phantoms.ct_body) at 0.08 cm
   voxels are each simulated once at a high history count (the target) and at
   several low counts (the inputs) with CT spiral beams of random pitch,
   voltage and tube count.  Inputs are the device features of the low run
   (Denoiser.world_features, the same path Denoiser.apply uses); the target is
   the high run's dose over the low run's D_ref; the loss mask is the body.
   64^3 patches are cropped from each pair.
2. Training: Adam on random batches of patches with flips along x, y and z.
3. Evaluation on the full CT spiral case: 512 x 512 x 400 at 0.08 cm, seeded
   body, 120 kVp dual source 2 x 40 mm, pitch 3.2.  A high-history run is the
   reference dose; plain runs and denoised runs at several history counts give
   the mean relative error over voxels above 20 % of the reference maximum;
   time to 2 % is interpolated on log-log axes from device time (transport
   kernels, plus features + inference + dose for the denoised runs).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import denoise as dn
from . import engine as dx
from . import phantoms
from .bowtie import DEFAULT_CT_BOWTIE, read_filters


def log(msg: str) -> None:
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def make_grid(ph: phantoms.Phantom, level: int = 1):
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=level))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    return world, grid


def spiral_beam(z_half: float, pitch: float, kv: float, dual: bool, histories: float, start_deg: float = 0.0):
    bow = read_filters()[DEFAULT_CT_BOWTIE]
    if dual:
        beam = dx.CTSpiralDualEnergyBeam((0.0, 0.0, -z_half), (0.0, 0.0, z_half), filtration={13: 6.0})
        beam.setTubeAVoltage(kv)
        beam.setTubeBVoltage(kv)
        beam.setBowtieFilterA(bow)
        beam.setBowtieFilterB(bow)
    else:
        beam = dx.CTSpiralBeam((0.0, 0.0, -z_half), (0.0, 0.0, z_half), filtration={13: 6.0})
        beam.setTubeVoltage(kv)
        beam.setBowtieFilter(bow)
    beam.setCollimation(4.0)
    beam.setPitch(pitch)
    beam.setStepAngleDeg(5.0)
    beam.setStartAngleDeg(start_deg)
    beam.setSourceDetectorDistance(119.0)
    beam.setNumberOfParticlesPerExposure(max(1, int(round(histories / beam.numberOfExposures()))))
    return beam


def run(world, grid, beam, seed: int) -> float:
    """One transport run; returns its device time in ms (transport kernels)."""
    dx.Transport(seed=seed)(world, beam, None, True)
    return float(grid.counters()["kernel_ms"])


# ------------------------------------------------------------------ training data

def generate(net: dn.Denoiser, n_phantoms: int, dims, hpv_target: float, hpv_inputs, patches: int, seed: int,
             phantom_seed0: int = 1000):
    """Patches (features, target, mask) from engine-generated pairs of the
    seeded phantoms phantom_seed0, phantom_seed0 + 1, ..."""
    rng = np.random.default_rng(seed)
    pool = []
    emb = None
    for i in range(n_phantoms):
        ph = phantoms.ct_body(dims, (0.08, 0.08, 0.08), seed=phantom_seed0 + i)
        if emb is None:
            emb = dn.material_embedding(ph.materials)
        world, grid = make_grid(ph)
        nx, ny, nz = ph.dims
        z_half = 0.5 * nz * 0.08
        pitch = float(rng.uniform(1.0, 3.5))
        kv = float(rng.choice([100.0, 120.0, 140.0]))
        dual = bool(rng.random() < 0.5)
        nv = ph.size
        run(world, grid, spiral_beam(z_half, pitch, kv, dual, hpv_target * nv), seed=10 * i + 1)
        # doses are totals over a run: compare them per history
        d_hi = grid.doseScored()[0].reshape(nz, ny, nx) / grid.histories()
        mat = ph.material.reshape(nz, ny, nx)
        mask = (mat != 0).astype(np.float32)
        for j, hpv in enumerate(hpv_inputs):
            grid.clearScores()
            run(world, grid, spiral_beam(z_half, pitch, kv, dual, hpv * nv), seed=10 * i + 2 + j)
            f = net.world_features(grid)[:, :, :nx]
            d_lo = grid.doseScored()[0].reshape(nz, ny, nx) / grid.histories()
            ref = dn.dose_reference(d_lo, mat)
            t = (d_hi / ref).astype(np.float32)
            for _ in range(patches):
                z0 = int(rng.integers(0, nz - 64 + 1))
                y0 = int(rng.integers(0, ny - 64 + 1))
                x0 = int(rng.integers(0, nx - 64 + 1))
                sl = (slice(z0, z0 + 64), slice(y0, y0 + 64), slice(x0, x0 + 64))
                if mask[sl].mean() < 0.05:
                    continue
                pool.append((f[sl].copy(), t[sl].copy(), mask[sl].copy()))
        log(f"phantom {i + 1}/{n_phantoms}: pitch {pitch:.2f}, {kv:.0f} kV, dual={dual}, {len(pool)} patches")
        del world, grid
    return pool


def batches(pool, batch: int, rng):
    while True:
        idx = rng.integers(0, len(pool), batch)
        f = np.stack([pool[k][0] for k in idx])
        t = np.stack([pool[k][1] for k in idx])
        m = np.stack([pool[k][2] for k in idx])
        for ax in (1, 2, 3):
            if rng.random() < 0.5:
                f = np.flip(f, axis=ax)
                t = np.flip(t, axis=ax)
                m = np.flip(m, axis=ax)
        yield f, t, m


def train(net: dn.Denoiser, pool, steps: int, batch: int, lr: float, seed: int):
    rng = np.random.default_rng(seed)
    gen = batches(pool, batch, rng)
    hist, ms = [], []
    for s in range(steps):
        # cosine decay to 10 % of the initial rate
        rate = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * s / max(1, steps))))
        f, t, m = next(gen)
        hist.append(net.train_step(f, t, m, lr=rate))
        ms.append(net.last_ms())
        if (s + 1) % 100 == 0 or s == 0:
            log(f"step {s + 1}/{steps}: loss {np.mean(hist[-100:]):.5f}, {np.mean(ms[-100:]):.1f} ms/step")
    return hist, ms


# ------------------------------------------------------------------ evaluation

def interpolate_crossing(hs, errs, target: float):
    """History count where the log-log interpolated error reaches target, or
    None when the measured range does not bracket it."""
    pts = sorted(zip(hs, errs))
    for (h0, e0), (h1, e1) in zip(pts, pts[1:]):
        if (e0 - target) * (e1 - target) <= 0 and e0 != e1:
            a = (math.log(target) - math.log(e0)) / (math.log(e1) - math.log(e0))
            return math.exp(math.log(h0) + a * (math.log(h1) - math.log(h0)))
    return None


def evaluate(net: dn.Denoiser, ref_histories: float, plain_hist, dn_hist, seed: int, target_err: float = 0.02,
             ph: phantoms.Phantom | None = None, log=log):
    """Time to target_err on the full CT spiral; `ph` is the phantom
    (default: the seeded 512 x 512 x 400 body), `log` the progress printer."""
    if ph is None:
        ph = phantoms.ct_body((512, 512, 400), (0.08, 0.08, 0.08), seed=seed)
    world, grid = make_grid(ph)
    z_half = 16.0

    def beam(h):
        return spiral_beam(z_half, 3.2, 120.0, True, h)

    # reference: accumulate runs of at most 2e10 histories
    grid.clearScores()
    done, ref_ms, k = 0.0, 0.0, 0
    while done < ref_histories:
        h = min(2e10, ref_histories - done)
        ref_ms += run(world, grid, beam(h), seed=9000 + k)
        done += h
        k += 1
        log(f"reference run {k}: {done:.2e} histories, {ref_ms / 1e3:.1f} s device")
    h_ref = grid.histories()
    d_ref, v_ref, _ = grid.doseScored()
    # the reference's worker deletes air dose by default: score body voxels only;
    # doses are totals over a run, so every comparison is per history
    d_ref = np.where(ph.material != 0, d_ref, 0.0) / h_ref
    v_ref = v_ref / float(h_ref) ** 2
    sel = d_ref > 0.2 * d_ref.max()
    r = d_ref[sel]
    ref_rel_var = v_ref[sel] / r ** 2
    ref_rel_se = float(np.sqrt(np.mean(ref_rel_var)))
    log(f"reference: {h_ref:.3e} histories, {ref_ms / 1e3:.1f} s, rms relative SE {ref_rel_se:.4f}, "
        f"{int(sel.sum())} voxels above 20 %")
    del v_ref

    def err(d, h):
        """(mean relative error, RMS relative error with the reference's own
        variance removed) of a run of h histories."""
        x = d[sel] / h
        rel = (x - r) / r
        mre = float(np.mean(np.abs(rel)))
        rms = float(np.sqrt(max(float(np.mean(rel * rel)) - float(np.mean(ref_rel_var)), 0.0)))
        return mre, rms

    plain = []
    for i, h in enumerate(plain_hist):
        grid.clearScores()
        ms = run(world, grid, beam(h), seed=100 + i)
        hh = grid.histories()
        d, v, _ = grid.doseScored()
        e, e_rms = err(d, hh)
        # the run's own variance estimate: Gaussian noise gives MRE = sqrt(2/pi) x relative SE
        own = float(np.sqrt(2.0 / np.pi) * np.mean(np.sqrt(np.maximum(v[sel], 0.0)) / np.maximum(d[sel], 1e-300)))
        plain.append({"histories": hh, "mc_ms": ms, "error": e, "error_rms_corrected": e_rms, "error_from_own_variance": own})
        log(f"plain {hh:.2e}: {ms:.1f} ms, error {e:.4f} (rms corrected {e_rms:.4f}, own-variance {own:.4f})")
    denoised = []
    for i, h in enumerate(dn_hist):
        grid.clearScores()
        ms = run(world, grid, beam(h), seed=200 + i)
        hh = grid.histories()
        e_in, _ = err(grid.doseScored()[0], hh)
        d = net.apply(grid)
        ams = net.last_ms()
        e, e_rms = err(d, hh)
        denoised.append({"histories": hh, "mc_ms": ms, "denoise_ms": ams, "error": e, "error_rms_corrected": e_rms,
                         "input_error": e_in})
        log(f"denoised {hh:.2e}: MC {ms:.1f} ms + denoise {ams:.1f} ms, error {e:.4f} (rms corrected {e_rms:.4f}, "
            f"input {e_in:.4f})")

    rate = sum(p["histories"] for p in plain) / (sum(p["mc_ms"] for p in plain) * 1e-3)
    # plain MC: error ~ a / sqrt(H), fitted on the measured points
    a = float(np.exp(np.mean([math.log(p["error"]) + 0.5 * math.log(p["histories"]) for p in plain])))
    h_plain = (a / target_err) ** 2
    t_plain = h_plain / rate
    h_dn = interpolate_crossing([p["histories"] for p in denoised], [p["error"] for p in denoised], target_err)
    t_dn = None
    if h_dn is not None:
        dn_ms = float(np.mean([p["denoise_ms"] for p in denoised]))
        t_dn = h_dn / rate + dn_ms * 1e-3
    return {
        "workload": "CT spiral 512x512x400 at 0.08 cm, seeded body, 120 kVp dual source 2x40 mm, pitch 3.2",
        "reference_histories": float(h_ref),
        "reference_rms_relative_se": ref_rel_se,
        "voxels_above_20pct": int(sel.sum()),
        "histories_per_s": rate,
        "plain": plain,
        "denoised": denoised,
        "plain_error_coefficient": a,
        "target_error": target_err,
        "plain_histories_to_target": h_plain,
        "plain_seconds_to_target": t_plain,
        "denoised_histories_to_target": h_dn,
        "denoised_seconds_to_target": t_dn,
        "speedup": (t_plain / t_dn) if t_dn else None,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--phantoms", type=int, default=12)
    ap.add_argument("--dims", type=int, nargs=3, default=(256, 256, 192), help="training phantom nx ny nz")
    ap.add_argument("--hpv-target", type=float, default=60.0, help="target histories per voxel")
    ap.add_argument("--hpv-inputs", type=float, nargs="+", default=(0.5, 1.0, 2.0, 4.0))
    ap.add_argument("--patches", type=int, default=6)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--load-weights", help="evaluate these weights (no training)")
    ap.add_argument("--init-weights", help="train starting from these weights")
    ap.add_argument("--phantom-seed0", type=int, default=1000, help="seed of the first training phantom")
    ap.add_argument("--save-weights")
    ap.add_argument("--skip-eval", action="store_true")
    ap.add_argument("--ref-histories", type=float, default=2e11)
    ap.add_argument("--plain", type=float, nargs="+", default=(2e9, 8e9))
    ap.add_argument("--denoised", type=float, nargs="+", default=(2.5e8, 5e8, 1e9, 2e9))
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)

    net = dn.Denoiser(seed=args.seed)
    result: dict = {"args": vars(args)}
    if args.load_weights:
        net.load(args.load_weights)
    else:
        if args.init_weights:
            net.load(args.init_weights)
        t0 = time.perf_counter()
        pool = generate(net, args.phantoms, tuple(args.dims), args.hpv_target, args.hpv_inputs, args.patches, args.seed,
                        args.phantom_seed0)
        result["data_s"] = time.perf_counter() - t0
        result["patches"] = len(pool)
        t0 = time.perf_counter()
        hist, ms = train(net, pool, args.steps, args.batch, args.lr, args.seed)
        result["train_s"] = time.perf_counter() - t0
        result["loss_first100"] = float(np.mean(hist[:100]))
        result["loss_last100"] = float(np.mean(hist[-100:]))
        result["train_ms_per_step"] = float(np.median(ms))
        del pool
        if args.save_weights:
            net.save(args.save_weights)
    if not args.skip_eval:
        result["evaluation"] = evaluate(net, args.ref_histories, args.plain, args.denoised, args.seed)
    with open(args.out, "w") as fh:
        json.dump(result, fh, indent=1)
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
