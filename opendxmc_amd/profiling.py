"""Profiling aids run under rocprofv3.

    python -m opendxmc_amd.profiling copy --mib 256 --iters 5

launches the library's plain device copy kernel; the FETCH_SIZE / WRITE_SIZE a
counter pass reads for it (expected: the copied bytes each way) gives the scale
factor applied to the transport kernel's counters.
"""
from __future__ import annotations

import argparse
import ctypes as C

from . import _native as N


def copy_probe(mib: int, iters: int) -> float:
    ms = C.c_double()
    N.check(N.lib().dxg_debug_device_copy(int(mib) << 20, int(iters), C.byref(ms)))
    return ms.value


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy"])
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--iters", type=int, default=5)
    a = ap.parse_args()
    ms = copy_probe(a.mib, a.iters)
    gbs = 2 * (a.mib << 20) / (ms * 1e-3) / 1e9
    print(f"copy {a.mib} MiB x {a.iters}: {ms:.3f} ms per launch, {gbs:.0f} GB/s (read+write)")


if __name__ == "__main__":
    main()
