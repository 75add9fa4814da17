"""Profiling aids run under rocprofv3.

    python -m opendxmc_amd.profiling copy --mib 256 --iters 5

launches the library's plain device copy kernel; the FETCH_SIZE / WRITE_SIZE a
counter pass reads for it (expected: the copied bytes each way) gives the scale
factor applied to the transport kernel's counters.

    python -m opendxmc_amd.profiling denoise --dims 400 512 512 --iters 3 --train-steps 5

times the denoiser: inference on a random volume (nz ny nx) and training steps
on a batch of four 64^3 patches; each convolution is 55,296 FLOP per voxel.

    python -m opendxmc_amd.profiling pipelines --dims 1024 512 512 --iters 5

times CT segmentation and the organ dose table on device-resident volumes
(17 bytes of HBM traffic per voxel each): organ indices in runs of 1..400
voxels (mean 200), then one organ drawn at random per voxel (the worst case).
"""
from __future__ import annotations

import argparse
import ctypes as C

import numpy as np

from . import _native as N

CONV_FLOP_PER_VOXEL = 2 * 27 * 32 * 32
LAYERS = 18


def copy_probe(mib: int, iters: int) -> float:
    ms = C.c_double()
    N.check(N.lib().dxg_debug_device_copy(int(mib) << 20, int(iters), C.byref(ms)))
    return ms.value


def denoise_probe(dims, iters: int, train_steps: int, batch: int = 4) -> None:
    from .denoise import Denoiser

    net = Denoiser(seed=1)
    rng = np.random.default_rng(1)
    f = rng.uniform(0.0, 2.0, tuple(dims) + (4,)).astype(np.float32)
    nv = int(np.prod(dims))
    for i in range(iters):
        net.infer(f)
        ms = net.last_ms()
        tflops = LAYERS * CONV_FLOP_PER_VOXEL * nv / (ms * 1e-3) / 1e12
        print(f"infer {tuple(dims)} #{i}: {ms:.2f} ms device, {tflops:.1f} TFLOP/s over the 18 convolutions",
              flush=True)
    del f
    if train_steps:
        pf = rng.uniform(0.0, 2.0, (batch, 64, 64, 64, 4)).astype(np.float32)
        t = rng.uniform(0.0, 2.0, (batch, 64, 64, 64)).astype(np.float32)
        m = np.ones_like(t)
        for i in range(train_steps):
            loss = net.train_step(pf, t, m, lr=1e-4)
            ms = net.last_ms()
            # forward 18 convs, backward 17 input-gradient + 18 weight-gradient convs
            flop = (18 + 17 + 18) * CONV_FLOP_PER_VOXEL * pf[..., 0].size
            print(f"train step #{i} (batch {batch} x 64^3): {ms:.2f} ms device, {flop / (ms * 1e-3) / 1e12:.1f} TFLOP/s,"
                  f" loss {loss:.4f}", flush=True)


def pipelines_probe(dims, iters: int) -> None:
    from . import pipelines as P
    from .devmem import DeviceBuffer

    n = int(np.prod(dims))
    rng = np.random.default_rng(1)
    tables = P.CTSegmentationPipeline().tables()
    hu = DeviceBuffer.from_array(rng.uniform(-1100.0, 2500.0, n))
    mat, dens = DeviceBuffer(n), DeviceBuffer(8 * n)
    for i in range(iters):
        P.segment_device(tables, hu.ptr, n, mat.ptr, dens.ptr)
        ms = P.last_ms(0)
        print(f"segment {tuple(dims)} #{i}: {ms:.3f} ms, {17 * n / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
    hu.free()
    dose = DeviceBuffer.from_array(rng.uniform(0.0, 1.0, n))
    runs = rng.integers(1, 400, n // 100 + 2)
    organs = {
        "runs": np.repeat((np.arange(runs.size) % 60).astype(np.uint8), runs)[:n],
        "random": rng.integers(0, 256, n, dtype=np.uint8),
    }
    sp = (0.1, 0.1, 0.1)
    for kind, organ in organs.items():
        mat.upload(organ)
        for i in range(iters):
            P.organ_dose_table_device(mat.ptr, dose.ptr, dens.ptr, n, sp)
            ms = P.last_ms(1)
            print(f"dose table {kind} {tuple(dims)} #{i}: {ms:.3f} ms, {17 * n / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "denoise", "pipelines"])
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--dims", type=int, nargs=3, default=(400, 512, 512))
    ap.add_argument("--train-steps", type=int, default=5)
    a = ap.parse_args()
    if a.what == "copy":
        ms = copy_probe(a.mib, a.iters)
        gbs = 2 * (a.mib << 20) / (ms * 1e-3) / 1e9
        print(f"copy {a.mib} MiB x {a.iters}: {ms:.3f} ms per launch, {gbs:.0f} GB/s (read+write)")
    elif a.what == "pipelines":
        pipelines_probe(a.dims, a.iters)
    else:
        denoise_probe(a.dims, a.iters, a.train_steps)


if __name__ == "__main__":
    main()
