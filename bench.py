"""Benchmark: histories/s of the voxel transport path on one MI355X.

Workload (one step = one Transport call over the whole beam):
  CT spiral, plain MC: 256 x 256 x 200 grid at 0.15 cm holding a seeded
  elliptic body (soft tissue, adipose shell, lungs 0.26 g/cm3, spine and ribs
  1.92 g/cm3, air), 120 kVp tungsten spectrum with 6 mm Al, the reference's
  default CT bowtie "Siemens Definition Flash W1 120kV", 4 cm collimation,
  pitch 1, 5 degree steps, SDD 119 cm, from z = -15 to +15 cm; 1e8 histories,
  FP32 transport, FP64 tally, correction level 1.
The grid, tables and tallies are resident in device memory before the timed
region (World::build() runs once, outside it).

Roofline: algorithmic bytes per history = 5 S + 24 D, with S the Woodcock
steps that read a voxel (1 B material + 4 B density gathered) and D the scored
deposits (24 B of FP64/u64 tally state), both counted on the device; achieved
= those bytes over the transport kernels' own duration (HIP events on the
stream they run on), against 8 TB/s HBM.  Macro-cell boundary crossings (one
read of a small cache-resident majorant array, no voxel read) are reported
beside them.

CPU baseline: the oracle (FP64 C++ restatement, std::thread) on a bounded
sample of the same workload, on this host's cores.

--workload ct-segmentation / organ-dose-table time the image pipelines on a
device-resident 1024 x 512 x 512 volume instead (voxels/s; 17 B of HBM traffic
per voxel each; CPU baseline: the oracle's numpy restatement, one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# the committed transport profiles with measured fabric traffic, per workload
TRAFFIC_PROFILES = {
    "ct-spiral-256": "profiles/transport_kernel_ct_spiral_1e8.md",
    "pencil-64": "profiles/transport_kernel_pencil_64_1e8.md",
    "ct-spiral-512-dual": "profiles/transport_kernel_ct_spiral_dual_512x512x400.md",
}


def measured_traffic_per_history(workload: str = "ct-spiral-256"):
    """Fabric traffic per history of a workload from its committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (copy-kernel calibrated); None when not
    profiled."""
    rel = TRAFFIC_PROFILES.get(workload)
    if rel is None:
        return None
    try:
        for line in (ROOT / rel).read_text().splitlines():
            if line.startswith("traffic_bytes_per_history:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_cpu() -> dict:
    """The host CPU: model, logical CPUs, the CPUs this process may run on,
    and the cgroup CPU quota when one is set; `threads` is what a CPU leg
    should use (the allowed CPUs, capped by the quota)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = logical
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    threads = allowed if quota is None else max(1, min(allowed, int(quota)))
    return {"model": model, "logical_cpus": logical, "allowed_cpus": allowed, "cgroup_cpu_quota": quota, "threads": threads}


def time_to_2pct(ref_histories: float, cpu_seconds: float | None):
    """Time to 2 % mean relative dose error on the full dual-source CT spiral
    (512 x 512 x 400 at 0.08 cm), measured in this run: a reference run of
    ref_histories, plain runs at 1e9, 4e9 and 1.6e10 histories (error a /
    sqrt(H) fitted, time at the measured rate) and runs of 2e7 .. 1.6e9
    histories denoised by the shipped network (2 % crossing interpolated on
    log-log axes; MC time plus the denoiser's features + inference + dose).
    With cpu_seconds, the oracle's rate on a sample of the same workload gives
    the CPU time for the plain histories to 2 %."""
    from opendxmc_amd import denoise as dn
    from opendxmc_amd import denoise_study as study
    from opendxmc_amd import phantoms

    def log(msg):
        print(f"[time-to-2%] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    ph = phantoms.ct_body((512, 512, 400), (0.08, 0.08, 0.08), seed=1)
    net = dn.Denoiser.pretrained()
    ev = study.evaluate(net, ref_histories, (1e9, 4e9, 1.6e10), (2e7, 3e7, 5e7, 1e8, 4e8, 1.6e9), seed=1, ph=ph, log=log)
    del net
    out = {
        "workload": "CT spiral dual source 512x512x400 at 0.08 cm, seeded body, 2 x 120 kVp 6 mm Al, W1 bowtie, "
                    "2 x 4 cm collimation, pitch 3.2, 32 cm; error = mean |D - D_ref| / D_ref over body voxels above "
                    "20 % of the reference maximum",
        "measured_in_this_run": True,
        "reference_histories": ev["reference_histories"],
        "reference_rms_relative_se": ev["reference_rms_relative_se"],
        "gpu_histories_per_s": ev["histories_per_s"],
        "plain_points": [(p["histories"], p["error"]) for p in ev["plain"]],
        "denoised_points": [(p["histories"], p["error"], p["denoise_ms"]) for p in ev["denoised"]],
        "plain_histories_to_2pct": ev["plain_histories_to_target"],
        "plain_gpu_mc_s": ev["plain_seconds_to_target"],
        "mc_plus_denoiser_histories_to_2pct": ev["denoised_histories_to_target"],
        "mc_plus_denoiser_s": ev["denoised_seconds_to_target"],
        "speedup_mc_plus_denoiser": ev["speedup"],
    }
    if cpu_seconds:
        _, beam = ct_dual_workload(1e9)
        cb = cpu_baseline(ph, beam, seconds_target=cpu_seconds, what="full dual-source CT spiral")
        t_cpu = ev["plain_histories_to_target"] / cb["value"]
        out["cpu_oracle_histories_per_s"] = cb["value"]
        out["cpu_oracle_s"] = t_cpu
        out["cpu_oracle_sample"] = cb["sample"]
        out["cpu_note"] = "CPU time = plain histories to 2 % / the oracle's measured rate (extrapolated)"
        out["speedup_plain_gpu_vs_cpu"] = t_cpu / ev["plain_seconds_to_target"]
        if ev["denoised_seconds_to_target"]:
            out["speedup_mc_plus_denoiser_vs_cpu"] = t_cpu / ev["denoised_seconds_to_target"]
    out["wall_s"] = time.perf_counter() - t0
    return out


def ct_workload(histories: int):
    import opendxmc_amd as dx
    from opendxmc_amd import phantoms
    from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters

    ph = phantoms.ct_body((256, 256, 200), (0.15, 0.15, 0.15), seed=1)
    beam = dx.CTSpiralBeam((0.0, 0.0, -15.0), (0.0, 0.0, 15.0), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setStepAngleDeg(5.0)
    beam.setSourceDetectorDistance(119.0)
    beam.setTubeVoltage(120.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    n_ex = beam.numberOfExposures()
    beam.setNumberOfParticlesPerExposure(max(1, round(histories / n_ex)))
    return ph, beam


def ct_dual_workload(histories: int):
    """The full CT spiral: 512 x 512 x 400 at 0.08 cm, seeded body, 120 kVp
    dual source, 2 x 4 cm collimation, pitch 3.2, 32 cm."""
    import opendxmc_amd as dx
    from opendxmc_amd import phantoms
    from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters

    ph = phantoms.ct_body((512, 512, 400), (0.08, 0.08, 0.08), seed=1)
    beam = dx.CTSpiralDualEnergyBeam((0.0, 0.0, -16.0), (0.0, 0.0, 16.0), {13: 6.0})
    beam.setTubeAVoltage(120.0)
    beam.setTubeBVoltage(120.0)
    beam.setCollimation(4.0)
    beam.setPitch(3.2)
    beam.setStepAngleDeg(5.0)
    beam.setSourceDetectorDistance(119.0)
    bow = read_filters()[DEFAULT_CT_BOWTIE]
    beam.setBowtieFilterA(bow)
    beam.setBowtieFilterB(bow)
    beam.setNumberOfParticlesPerExposure(max(1, round(histories / beam.numberOfExposures())))
    return ph, beam


def pencil_workload(histories: int):
    """Plain MC, small: a 60 keV pencil beam along +z (through a voxel centre,
    x = y = 0.05 cm) into a 64^3 water cube at 0.1 cm."""
    import opendxmc_amd as dx
    from opendxmc_amd import phantoms

    ph = phantoms.water_cube(64, 0.1)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0.0, 0.0, 1.0), 60.0)
    beam.setNumberOfParticlesPerExposure(int(histories))
    return ph, beam


WORKLOADS = {
    "pencil-64": (pencil_workload, 1e6, "Plain MC small, 60 keV pencil beam into a 64^3 water cube at 0.1 cm"),
    "ct-spiral-256": (ct_workload, 1e8, "CT spiral plain MC, 256x256x200 at 0.15 cm seeded body phantom, 120 kVp 6 mm Al, "
                                        "Siemens Definition Flash W1 120kV bowtie, 4 cm collimation, pitch 1, 30 cm"),
    "ct-spiral-512-dual": (ct_dual_workload, 1e9, "CT spiral dual source plain MC, 512x512x400 at 0.08 cm seeded body "
                                                  "phantom, 2 x 120 kVp 6 mm Al, W1 bowtie, 2 x 4 cm collimation, "
                                                  "pitch 3.2, 32 cm"),
}


PIPE_PROFILE = ROOT / "profiles" / "ct_segmentation_and_organ_dose_table_1024x512x512.md"
PIPE_DIMS = (1024, 512, 512)
PIPE_KERNELS = {"ct-segmentation": "segment_pairs_kernel", "organ-dose-table": "organ_table_kernel"}


def pipeline_traffic(kernel: str):
    """HBM bytes per launch of a pipeline kernel from the committed counter
    passes; None when not profiled."""
    try:
        for line in PIPE_PROFILE.read_text().splitlines():
            if line.startswith(f"traffic_bytes_per_launch_{kernel}:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def pipeline_inputs(workload: str, n: int):
    rng = np.random.default_rng(1)
    if workload == "ct-segmentation":
        return {"hu": rng.uniform(-1100.0, 2500.0, n)}
    runs = rng.integers(1, 400, n // 100 + 2)
    return {"organ": np.repeat((np.arange(runs.size) % 60).astype(np.uint8), runs)[:n],
            "dose": rng.uniform(0.0, 1.0, n), "density": rng.uniform(0.0, 2.0, n)}


def pipeline_cpu_baseline(workload: str, tables, seconds_target: float = 10.0):
    """The oracle's numpy restatement, one thread, repeated over a 2^24-voxel
    sample of the same inputs for about seconds_target."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    n = 1 << 24
    x = pipeline_inputs(workload, n)
    sep = np.array(tables.separators[: tables.n_materials - 1])
    att = np.array(tables.attenuation[: tables.n_materials])
    done, t0 = 0, time.perf_counter()
    while True:
        if workload == "ct-segmentation":
            oracle.ct_segment(sep, att, tables.slope, tables.offset, x["hu"])
        else:
            oracle.organ_dose_table(x["organ"], x["dose"], x["density"], (0.1, 0.1, 0.1))
        done += n
        dt = time.perf_counter() - t0
        if dt >= seconds_target:
            break
    return {"value": done / dt, "unit": "voxels/s", "cores": 1, "kind": "port",
            "sample": f"{done // n} passes over 2^24 voxels of the same inputs, numpy oracle, 1 thread, {dt:.1f} s"}


def pipeline_bench(args) -> int:
    """One step = one pass of the segmentation or dose-table kernel over a
    device-resident 1024 x 512 x 512 volume."""
    from opendxmc_amd import pipelines as P
    from opendxmc_amd.devmem import DeviceBuffer

    n = int(np.prod(PIPE_DIMS))
    tables = P.CTSegmentationPipeline().tables()
    x = pipeline_inputs(args.workload, n)
    if args.workload == "ct-segmentation":
        bufs = [DeviceBuffer.from_array(x["hu"]), DeviceBuffer(n), DeviceBuffer(8 * n)]

        def step():
            P.segment_device(tables, bufs[0].ptr, n, bufs[1].ptr, bufs[2].ptr)
        which, dtype = 0, "f64"
        description = ("CT number -> material / density segmentation (ctsegmentationpipeline.cpp), 5 materials at "
                       "120 kV 0.9 mm Al, CT numbers uniform in [-1100, 2500)")
    else:
        bufs = [DeviceBuffer.from_array(x["organ"]), DeviceBuffer.from_array(x["dose"]),
                DeviceBuffer.from_array(x["density"])]

        def step():
            P.organ_dose_table_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n, (0.1, 0.1, 0.1))
        which, dtype = 1, "f64"
        description = ("organ dose table (dosetablepipeline.cpp), 60 organs in runs of 1..400 voxels, FP64 dose "
                       "and density")
    del x
    for _ in range(args.warmup):
        step()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kernel_ms += P.last_ms(which)
    wall = time.perf_counter() - t0
    per_launch_ms = kernel_ms / max(args.steps, 1)
    achieved = 17.0 * n / (per_launch_ms * 1e-3) / 1e9
    result = {
        "metric": "voxels/s",
        "value": n * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"{description}, {PIPE_DIMS[0]}x{PIPE_DIMS[1]}x{PIPE_DIMS[2]} voxels resident in HBM"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pipeline_traffic(PIPE_KERNELS[args.workload]),
            "traffic_source": str(PIPE_PROFILE.relative_to(ROOT)) + " (FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch)",
            "kernel": PIPE_KERNELS[args.workload],
            "kernel_ms_per_launch": per_launch_ms,
            "bytes_per_voxel": 17,
        },
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = pipeline_cpu_baseline(args.workload, tables)
    print(json.dumps(result), flush=True)
    return 0


def cpu_baseline(ph, beam, seconds_target: float = 15.0, what: str = "CT spiral"):
    """Time the oracle on a bounded sample: every exposure, fewer particles,
    on every CPU this process may use (within the cgroup quota)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    host = host_cpu()
    threads = host["threads"]
    ex = beam.exposures()
    spectra = beam.spectra()
    bows = [(b.angle, b.weight) for b in beam.bowties()]
    # probe with a tiny sample, then size the timed sample to ~seconds_target
    probe = ex.copy()
    probe["n_particles"] = 200
    t0 = time.perf_counter()
    oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, ph.materials, probe, spectra, bows,
                     level=1, seed=11, threads=threads)
    rate = probe["n_particles"].sum() / max(time.perf_counter() - t0, 1e-6)
    # the probe's rate includes thread start-up, so a sample sized from it can
    # run well short of the target; one resize from the measured rate fixes that
    cap = int(ex["n_particles"][0])
    for seed in (12, 13):
        per_ex = int(max(50, min(rate * seconds_target / len(ex), cap)))
        sample = ex.copy()
        sample["n_particles"] = per_ex
        t0 = time.perf_counter()
        _, _, _, c = oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, ph.materials, sample, spectra,
                                      bows, level=1, seed=seed, threads=threads)
        dt = time.perf_counter() - t0
        rate = c["histories"] / dt
        if dt >= 0.5 * seconds_target or per_ex >= cap:
            break
    return {
        "value": c["histories"] / dt,
        "unit": "histories/s",
        "cores": threads,
        "kind": "port",
        "host": host,
        "sample": f"{c['histories']} histories ({per_ex} per exposure x {len(ex)} exposures) of the same {what} "
                  f"workload, FP64 oracle, {threads} threads on {host['model']} ({host['logical_cpus']} logical CPUs, "
                  f"{host['allowed_cpus']} allowed, quota {host['cgroup_cpu_quota']}), {dt:.1f} s",
        "steps_per_history": c["steps"] / c["histories"],
        "deposits_per_history": c["deposits"] / c["histories"],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + sorted(PIPE_KERNELS), default="ct-spiral-256")
    ap.add_argument("--histories", type=float, default=None, help="histories per step (default per workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-2pct", action="store_true", help="skip the time-to-2 %% measurement on the full CT spiral")
    ap.add_argument("--t2-ref-histories", type=float, default=3e11, help="reference histories of the time-to-2 %% leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (histories shard trivially "
              f"across GPUs, but this build measures one MI355X)", file=sys.stderr)
        return 2

    if args.workload in PIPE_KERNELS:
        return pipeline_bench(args)

    import opendxmc_amd as dx

    make, default_hist, description = WORKLOADS[args.workload]
    histories = int(args.histories if args.histories else default_hist)
    ph, beam = make(histories)
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    t0 = time.perf_counter()
    world.build()
    build_s = time.perf_counter() - t0
    transport = dx.Transport(seed=1)

    for i in range(args.warmup):
        transport.seed = 1000 + i
        transport(world, beam, None, True)
        print(f"warmup {i} done", file=sys.stderr, flush=True)
    c = grid.counters()   # the launch settings the warm-up runs chose
    launch = {"refill_threshold": c["refill_threshold"], "interact_threshold": c["interact_threshold"],
              "outer_threshold": c["outer_threshold"], "batch": c["batch"], "autotuned": bool(c["autotuned"])}
    steps = bytes_alg = kernel_ms = launches = hist = deposits = crossings = 0
    phase = [0, 0, 0, 0]
    phase_lanes = [0, 0, 0, 0]
    phase_iters = [0, 0, 0, 0]
    t0 = time.perf_counter()
    for i in range(args.steps):
        transport.seed = i + 1
        transport(world, beam, None, True)
        c = grid.counters()
        steps += c["steps"]
        deposits += c["deposits"]
        crossings += c["cell_crossings"]
        kernel_ms += c["kernel_ms"]
        launches += c["kernel_launches"]
        hist += c["histories"]
        phase = [a + b for a, b in zip(phase, c["phase_cycles"])]
        phase_lanes = [a + b for a, b in zip(phase_lanes, c["phase_lanes"])]
        phase_iters = [a + b for a, b in zip(phase_iters, c["phase_iters"])]
        print(f"step {i}: {c['histories']} histories, kernel {c['kernel_ms']:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    bytes_alg = 5.0 * steps + 24.0 * deposits
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    tph = measured_traffic_per_history(args.workload)
    hist_per_launch = hist / max(launches, 1)
    traffic = tph * hist_per_launch if tph is not None else None
    result = {
        "metric": "histories/s",
        "value": hist / wall,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{description}, {histories:.0e} histories per step, correction level 1, FP64 tally",
            "histories_per_step": hist // max(args.steps, 1),
            "exposures": beam.numberOfExposures(),
            "world_build_s": build_s,
            "steps_per_history": steps / max(hist, 1),
            "deposits_per_history": deposits / max(hist, 1),
            "cell_crossings_per_history": crossings / max(hist, 1),
            "kernel_launches_per_step": launches / max(args.steps, 1),
            "launch_settings": launch,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"{TRAFFIC_PROFILES[args.workload]} (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, per history, "
                               "times histories per launch)") if traffic is not None else None,
            "kernel": "transport_kernel",
            "kernel_ms_per_launch": kernel_ms / max(launches, 1),
            "bytes_per_history": bytes_alg / max(hist, 1),
        },
    }
    if sum(phase) > 0:   # libraries built with DXG_PHASE_CLOCKS
        names = ("refill", "interact", "outer", "step")
        result["config"]["phase_cycle_share"] = {k: v / sum(phase) for k, v in zip(names, phase)}
        result["config"]["phase_lane_use"] = {k: ln / (64.0 * it) if it else 0.0 for k, ln, it in zip(names, phase_lanes, phase_iters)}
        result["config"]["phase_iters_per_history"] = {k: it / max(hist, 1) for k, it in zip(names, phase_iters)}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ph, beam, what=description)
    if not args.no_time_to_2pct and args.workload == "ct-spiral-256":
        del world, grid
        result["time_to_2pct"] = time_to_2pct(args.t2_ref_histories, None if args.no_cpu_baseline else 10.0)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
