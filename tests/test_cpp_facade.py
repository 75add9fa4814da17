"""The C++ facade (include/dxmc/) with DXMClib's call surface, exercised by a
headless reproduction of OpenDXMC's worker<C> (tests/cpp/worker_harness.cpp,
after simulationpipeline.cpp:124-235)."""
from __future__ import annotations

import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import build, phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters


def harness() -> Path:
    if not build.HARNESS.exists():
        build.build_harness()
    return build.HARNESS


def write_phantom(path: Path, ph: phantoms.Phantom) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<3Q", *ph.dims))
        f.write(struct.pack("<3d", *ph.spacing))
        f.write(struct.pack("<Q", len(ph.materials)))
        for m in ph.materials:
            f.write(struct.pack("<Q", len(m)))
            for z, w in m.items():
                f.write(struct.pack("<Qd", z, w))
        f.write(np.ascontiguousarray(ph.density, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(ph.material, dtype=np.uint8).tobytes())


def write_bowtie(path: Path) -> dx.BowtieFilter:
    b = read_filters()[DEFAULT_CT_BOWTIE]
    path.write_text("\n".join(f"{float(a)!r} {float(w)!r}" for a, w in zip(b.angle, b.weight)))
    assert len(path.read_text().split()) == 2 * len(b.angle)
    return b


def test_harness_builds_and_reports_failure_without_device(tmp_path):
    exe = harness()
    assert exe.exists()
    import os
    if Path("/dev/kfd").exists():
        pytest.skip("a GPU is visible here")
    ph = phantoms.water_cube(4, 1.0)
    write_phantom(tmp_path / "p.bin", ph)
    r = subprocess.run([str(exe), str(tmp_path / "p.bin"), "-", str(tmp_path / "o.bin"), "1", "1", "100"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "simulation failed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("ctdivol", [None, 1.0])
def test_worker_harness_matches_python_path(tmp_path, ctdivol):
    """worker<C> on the C++ facade and the Python beams through one C-ABI give
    the same doses, both with default beams (no calibration value: uncalibrated
    in both facades) and with a CTDIvol of 1 mGy set on both."""
    ph = phantoms.ct_body((40, 40, 24), (0.9, 0.9, 0.9), seed=8)
    write_phantom(tmp_path / "p.bin", ph)
    bow = write_bowtie(tmp_path / "b.txt")
    seed, ppe = 77, 20000
    print(f"[harness] dims={ph.dims} seed={seed} ppe={ppe} ctdivol={ctdivol}", flush=True)
    extra = [] if ctdivol is None else [str(ctdivol)]
    r = subprocess.run([str(harness()), str(tmp_path / "p.bin"), str(tmp_path / "b.txt"), str(tmp_path / "o.bin"),
                        "1", str(seed), str(ppe), *extra], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    n = ph.size
    raw = (tmp_path / "o.bin").read_bytes()
    dose = np.frombuffer(raw[: 8 * n], dtype="<f8")
    count = np.frombuffer(raw[8 * n: 16 * n], dtype="<f8")
    var = np.frombuffer(raw[16 * n: 24 * n], dtype="<f8")
    units = raw[24 * n:].decode()

    # the same run through the Python mirror of the same C-ABI
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    hz = 0.5 * ph.dims[2] * ph.spacing[2]
    beam = dx.CTSpiralBeam((0, 0, -hz), (0, 0, hz), {13: 6.0})
    beam.setTubeVoltage(120)
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setStepAngleDeg(15)
    beam.setBowtieFilter(bow)
    beam.setNumberOfParticlesPerExposure(ppe)
    # worker<C> runs with useBeamCalibration (simulationpipeline.cpp:165)
    if ctdivol is not None:
        beam.setCTDIvol(ctdivol)
        beam.setCTDIdiameter(32.0)
    assert dx.Transport(seed=seed)(world, beam, None, True)
    d, v, c = grid.doseScored()
    air = ph.material == 0
    d[air] = 0.0
    v[air] = 0.0
    c = c.astype(float)
    c[air] = 0.0
    scale = 1e3 if d.max() < 1 else 1.0
    assert units == ("uGy" if scale == 1e3 else "mGy")
    np.testing.assert_allclose(dose, d * scale, rtol=1e-9, atol=1e-300)
    np.testing.assert_allclose(var, v * scale * scale, rtol=1e-6, atol=1e-300)
    np.testing.assert_array_equal(count, c)
    assert np.all(dose[air] == 0) and dose.max() > 0
