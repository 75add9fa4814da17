"""Build the native parts in-tree.

* ``opendxmc_amd/libdxmc_hip.so``: HIP kernels for gfx950 plus the C-ABI
  (include/dxmc_hip.h), compiled with hipcc.
* ``opendxmc_amd/physics/physics_list.bin``: the physics list, regenerated
  when its generator changed.
* ``oracle/liboracle_dxmc.so``: the CPU oracle (test infrastructure), g++.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "opendxmc_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libdxmc_hip.so"
PHYS = PKG / "physics" / "physics_list.bin"
PHYS_GEN = PKG / "physics" / "gen_physics_list.py"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle_dxmc.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DXG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
SOURCES = ["physics.cpp", "world.cpp", "beams.cpp", "calibration.cpp", "batch.cpp", "transport.hip", "denoise.hip", "pipelines.hip",
           "h5io.cpp"]
# h5io.cpp compiles against the HDF5 headers only; the library itself is
# opened with dlopen when a project file is first touched.
HDF5_INCLUDE = Path(os.environ.get("DXG_HDF5_INCLUDE", "/opt/conda/include"))
EXTRA_FLAGS = {"h5io.cpp": ["-isystem", str(HDF5_INCLUDE)]}


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed: {' '.join(cmd)}")


def build_physics_list(force: bool = False) -> Path:
    stamp = PHYS.with_suffix(".sha256")
    digest = hashlib.sha256(PHYS_GEN.read_bytes()).hexdigest()
    if not force and PHYS.exists() and stamp.exists() and stamp.read_text().strip() == digest:
        return PHYS
    _run([sys.executable, str(PHYS_GEN), str(PHYS)])
    stamp.write_text(digest + "\n")
    return PHYS


def build_library() -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = BUILD / (src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append(([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), *lang, "-c", str(path), "-o", str(obj)], str(obj)))
    # one compiler per source file, at most one per CPU
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 1))) as pool:
        list(pool.map(_run, [cmd for cmd, _ in jobs]))
    objs = [obj for _, obj in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp), "-ldl", "-lz", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR), "-B", "liboracle_dxmc.so"])
    return ORACLE_LIB


HARNESS_SRC = ROOT / "tests" / "cpp" / "worker_harness.cpp"
HARNESS = ROOT / "tests" / "cpp" / "_build" / "worker_harness"


def build_harness() -> Path:
    """The C++ reproduction of the reference's worker<C>, on the C++ facade."""
    HARNESS.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(HARNESS_SRC),
          f"-L{PKG}", "-ldxmc_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../../opendxmc_amd",
          "-o", str(HARNESS)])
    return HARNESS


def build_all() -> None:
    build_physics_list()
    build_library()
    build_oracle()
    build_harness()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}, {PHYS}, {ORACLE_LIB}")
