"""The physics list's known photoelectric deficit in dose terms.

The list's photoelectric cross sections come from screened-hydrogenic models
and run 5-15 % low below 40 keV against NIST totals (checked on CPU in
test_physics_list.py).  This test bounds what that does to dose: the same
runs with every element's photoelectric (and its K-shell part) raised by
15 % below 40 keV -- the largest deficit -- through a perturbed copy of the
list (DXG_PHYSICS_LIST, read once per process, so each list runs in a child
process), for a 120 kVp CT spiral and a 70 kVp DX field on a seeded body.
Measured: +1.36 % (CT) and +3.15 % (DX) mean body dose; the test fails if
they grow past 3 % and 6 % or change sign."""
from __future__ import annotations

import json
import os
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
LIST = ROOT / "opendxmc_amd" / "physics" / "physics_list.bin"


def perturbed_list(src: Path, dst: Path, factor: float, below_kev: float) -> None:
    """A copy of the physics list with pe and pe_k scaled by `factor` at
    energies below `below_kev` (format: opendxmc_amd/physics/gen_physics_list.py)."""
    b = src.read_bytes()
    out = bytearray(b)
    o = 8
    zmax, ne, nx = struct.unpack_from("<III", b, o)
    o += 12 + 8 * ne + 8 * nx
    for _ in range(zmax):
        o += 8 + 16 * ne + 16 * nx            # A, incoh, coh, F, S
        (npe,) = struct.unpack_from("<I", b, o)
        o += 4
        e = np.frombuffer(b, "<f8", npe, o)
        scale = np.where(e < below_kev, factor, 1.0)
        for k in (1, 2):                        # pe, then pe_k
            at = o + 8 * npe * k
            v = np.frombuffer(b, "<f8", npe, at) * scale
            out[at:at + 8 * npe] = v.astype("<f8").tobytes()
        o += 24 * npe
        (nsh,) = struct.unpack_from("<I", b, o)
        o += 4 + 48 * nsh + 56
    assert o == len(b)
    dst.write_bytes(bytes(out))


CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import opendxmc_amd as dx
from opendxmc_amd import phantoms
ph = phantoms.ct_body((128, 128, 80), (0.3, 0.3, 0.3), seed=11)
world = dx.World()
grid = world.addItem(dx.AAVoxelGrid(correction=1))
grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
grid.setSpacing(ph.spacing)
world.build()
body = ph.material != 0
out = {}
ct = dx.CTSpiralBeam((0, 0, -10), (0, 0, 10), {13: 6.0})
ct.setCollimation(4.0)
ct.setPitch(1.0)
ct.setStepAngleDeg(10.0)
ct.setNumberOfParticlesPerExposure(40_000)
dxb = dx.DXBeam({13: 2.5})
dxb.tube.voltage = 70.0
dxb.setCollimation((25.0, 25.0))
dxb.setNumberOfParticlesPerExposure(6_000_000)
for name, beam in (("ct", ct), ("dx", dxb)):
    grid.clearScores()
    dx.Transport(seed=3)(world, beam, None, False)
    d = grid.doseScored()[0]
    out[name] = float(d[body].mean()) / grid.histories()
print(json.dumps(out))
"""


def _mean_doses(list_path: Path) -> dict:
    env = dict(os.environ, DXG_PHYSICS_LIST=str(list_path))
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT)], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_photoelectric_deficit_in_dose_terms(tmp_path):
    hi = tmp_path / "physics_pe_plus15.bin"
    perturbed_list(LIST, hi, 1.15, 40.0)
    assert hi.stat().st_size == LIST.stat().st_size
    base = _mean_doses(LIST)
    pert = _mean_doses(hi)
    change = {k: pert[k] / base[k] - 1.0 for k in base}
    print(f"[physics sensitivity] mean body dose change with photoelectric x1.15 below 40 keV: "
          f"CT 120 kVp {100 * change['ct']:+.2f} %, DX 70 kVp {100 * change['dx']:+.2f} %", flush=True)
    # more absorption below 40 keV: more dose in the body, more for the softer beam
    assert 0.0 < change["ct"] < 0.03
    assert 0.0 < change["dx"] < 0.06
