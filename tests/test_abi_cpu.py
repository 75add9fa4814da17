"""C-ABI library on the CPU: it loads, exports every entry point include/*.h
declares, rejects bad input the way the reference's readiness checks do
(simulationpipeline.cpp:45-63, 133-145), and its host-side parts -- material
attenuation, tube spectra, exposure geometry -- hold their invariants.  No
transport runs here (there is no GPU); tests/test_gpu_parity.py covers that.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import _native as N
from opendxmc_amd import phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(dxg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in N.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_correction_levels():
    with pytest.raises(N.NativeError) as e:
        dx.AAVoxelGrid(correction=7)
    assert e.value.code == N.DXG_ERR_ARGUMENT
    for level in (0, 1, 2):
        dx.AAVoxelGrid(correction=level)


def test_material_index_readiness_check():
    g = dx.AAVoxelGrid()
    mats = [dx.Material.byWeight(phantoms.AIR), dx.Material.byWeight(phantoms.WATER)]
    mat = np.ones(8, dtype=np.uint8)
    mat[3] = 2   # >= number of materials
    with pytest.raises(N.NativeError) as e:
        g.setData((2, 2, 2), np.ones(8), mat, mats)
    assert e.value.code == N.DXG_ERR_DATA
    with pytest.raises(N.NativeError) as e:
        g.setData((2, 2, 2), -np.ones(8), np.ones(8, dtype=np.uint8), mats)
    assert e.value.code == N.DXG_ERR_DATA


def test_material_by_weight_failure_is_none():
    assert dx.Material.byWeight({200: 1.0}) is None
    assert dx.Material.byWeight({8: 0.0}) is None
    assert dx.Material.byWeight({8: -1.0, 1: 2.0}) is None
    assert dx.Material.byWeight({1: 2.0, 8: 16.0}) is not None   # weights are normalised


def test_call_order_is_checked():
    g = dx.AAVoxelGrid()
    with pytest.raises(N.NativeError) as e:
        g.build()
    assert e.value.code == N.DXG_ERR_STATE
    w = dx.World()
    w.addItem(g)
    beam = dx.PencilBeam((0, 0, -10), (0, 0, 1), 60)
    with pytest.raises(N.NativeError) as e:
        dx.Transport()(w, beam)
    assert e.value.code == N.DXG_ERR_STATE
    with pytest.raises(N.NativeError):
        g.setSpacing((0.1, -0.1, 0.1))


def test_build_without_device_fails_loudly():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and Path("/dev/kfd").exists():
        pytest.skip("a GPU is visible here")
    ph = phantoms.water_cube(4, 1.0)
    g = dx.AAVoxelGrid()
    g.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    g.setSpacing(ph.spacing)
    with pytest.raises(N.NativeError) as e:
        g.build()
    assert e.value.code == N.DXG_ERR_DEVICE


def test_progress_object():
    p = dx.TransportProgress()
    assert p.continueSimulation()
    assert p.progress() == (0, 0)
    assert p.message() == ""
    p.setStopSimulation()
    assert not p.continueSimulation()


def test_missing_physics_list_is_reported():
    code = ("import opendxmc_amd as dx\n"
            "from opendxmc_amd import _native as N\n"
            "try:\n    dx.Material({8: 1.0}).attenuationValues([60.0])\n"
            "except N.NativeError as e:\n    print('CODE', e.code)\n")
    env = dict(os.environ, DXG_PHYSICS_LIST="/nonexistent/physics_list.bin", PYTHONPATH=str(ROOT))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert f"CODE {N.DXG_ERR_PHYSICS}" in out.stdout, out.stdout + out.stderr


@pytest.mark.parametrize("comp", [phantoms.WATER, phantoms.BONE, phantoms.LUNG, {53: 0.5, 1: 0.5}, {74: 1.0}])
def test_product_attenuation_matches_oracle(oracle, comp):
    e = np.concatenate([np.geomspace(1.0, 199.0, 57), [33.17, 33.2, 69.6]])
    prod = dx.Material(comp).attenuationValues(e)
    orc = oracle.attenuation(comp, e)
    np.testing.assert_allclose(prod, orc, rtol=1e-10)


def test_tube_spectrum_properties():
    for kv in (60.0, 80.0, 120.0, 140.0):
        e, w = dx.Tube(kv, 12.0, {13: 3.0}).getSpecter()
        assert w.sum() == pytest.approx(1.0, rel=1e-12)
        assert e.max() < kv and e.min() >= 1.0 and np.all(w >= 0)
        lines = [57.982, 59.318, 67.244, 69.067]
        has_lines = all(np.any(np.isclose(e, l)) for l in lines)
        assert has_lines == (kv > 69.525)
    # more filtration hardens the beam
    hvl = [dx.Tube(120.0, 12.0, {13: t}).halfValueLayerAl() for t in (2.0, 6.0, 9.0)]
    mean = [dx.Tube(120.0, 12.0, {13: t}).meanSpecterEnergy() for t in (2.0, 6.0, 9.0)]
    assert hvl[0] < hvl[1] < hvl[2] and mean[0] < mean[1] < mean[2]
    # higher voltage hardens it too
    assert dx.Tube(80.0, 12.0, {13: 3.0}).halfValueLayerAl() < dx.Tube(120.0, 12.0, {13: 3.0}).halfValueLayerAl()
    # a steeper anode (smaller angle) self-filters more
    assert dx.Tube(120.0, 7.0, {13: 3.0}).meanSpecterEnergy() > dx.Tube(120.0, 16.0, {13: 3.0}).meanSpecterEnergy()
    with pytest.raises(N.NativeError):
        dx.Tube(120.0, 12.0, {99: 1.0}).getSpecter()


def axis_of(ex):
    c0, c1 = ex["cosines"][:, :3], ex["cosines"][:, 3:]
    return c0, c1, np.cross(c0, c1)


def test_ct_spiral_geometry():
    b = dx.CTSpiralBeam((0, 0, -15), (0, 0, 15))
    b.setCollimation(4.0)
    b.setPitch(1.0)
    b.setStepAngleDeg(5.0)
    b.setSourceDetectorDistance(119.0)
    ex = b.exposures()
    assert len(ex) == int(30 / 4.0 * 72) + 1
    c0, c1, ax = axis_of(ex)
    np.testing.assert_allclose(np.linalg.norm(c0, axis=1), 1, atol=1e-12)
    np.testing.assert_allclose(np.sum(c0 * c1, axis=1), 0, atol=1e-12)
    pos = ex["position"]
    # sources on a helix of radius SDD/2, the beam axis through the rotation axis
    np.testing.assert_allclose(np.hypot(pos[:, 0], pos[:, 1]), 59.5, rtol=1e-12)
    t = -(pos[:, 0] * ax[:, 0] + pos[:, 1] * ax[:, 1])
    np.testing.assert_allclose(pos[:, :2] + t[:, None] * ax[:, :2], 0, atol=1e-9)
    # table advance per turn = pitch x collimation, gantry step 5 degrees
    np.testing.assert_allclose(pos[72, 2] - pos[0, 2], 4.0, rtol=1e-12)
    ang = np.arctan2(-pos[:, 1], -pos[:, 0])
    np.testing.assert_allclose(np.diff(np.unwrap(ang)), math.radians(5), atol=1e-9)
    assert ex["collimation_half_angles"][0, 1] == pytest.approx(math.atan(4.0 / 119.0))
    assert ex["collimation_half_angles"][0, 0] == pytest.approx(math.asin(50.0 / 119.0))


def test_ct_spiral_aec_and_dual_source():
    b = dx.CTSpiralDualEnergyBeam((0, 0, -10), (0, 0, 10))
    b.setCollimation(4.0)
    b.setPitch(2.0)
    b.setTubeBoffsetAngleDeg(95.0)
    b.setRelativeMasTubeB(0.5)
    b.setAECFilter(dx.CTAECFilter((0, 0, -10), (0, 0, 10), (1.0, 3.0)))
    ex = b.exposures()
    a, bb = ex[0::2], ex[1::2]
    assert np.all(a["spectrum_index"] == 0) and np.all(bb["spectrum_index"] == 1)
    ang_a = np.arctan2(-a["position"][:, 1], -a["position"][:, 0])
    ang_b = np.arctan2(-bb["position"][:, 1], -bb["position"][:, 0])
    np.testing.assert_allclose(np.mod(ang_b - ang_a, 2 * np.pi), math.radians(95.0), atol=1e-9)
    # AEC normalised to mean 1 along start->stop: 0.5 at start, 1.5 at stop
    assert a["weight"][0] == pytest.approx(0.5)
    assert a["weight"][-1] == pytest.approx(1.5, rel=1e-3)
    np.testing.assert_allclose(bb["weight"], 0.5 * a["weight"])


def test_sequential_and_cbct_geometry():
    s = dx.CTSequentialBeam((0, 0, -2.0), (0, 0, 1))
    s.setNumberOfSlices(3)
    s.setSliceSpacing(2.0)
    ex = s.exposures()
    assert len(ex) == 3 * 72
    np.testing.assert_allclose(np.unique(np.round(ex["position"][:, 2], 9)), [-2.0, 0.0, 2.0])
    c = dx.CBCTBeam((1.0, 2.0, 3.0), (0, 0, 1))
    c.setStartAngleDeg(-90)
    c.setStopAngleDeg(90)
    c.setStepAngleDeg(1)
    ex = c.exposures()
    assert len(ex) == 181
    _, _, ax = axis_of(ex)
    d = ex["position"] + 60.0 * ax - np.array([1.0, 2.0, 3.0])
    np.testing.assert_allclose(d, 0, atol=1e-9)


def test_dx_beam_matches_reference_geometry():
    # dxmc_specialization.cpp:21-90: defaults SPD 100, SDD 100, collimation 20 x 20 cm
    b = dx.DXBeam()
    ex = b.exposures()[0]
    np.testing.assert_allclose(ex["position"], (0.0, 100.0, 0.0), atol=1e-12)
    np.testing.assert_allclose(ex["cosines"], (0, 0, 1, -1, 0, 0), atol=1e-12)
    np.testing.assert_allclose(ex["collimation_half_angles"], math.tan(0.1))


def test_pencil_exposure():
    ex = dx.PencilBeam((1, 2, 3), (0, 0, 2), 60).exposures()[0]
    _, _, ax = axis_of(ex[None])
    np.testing.assert_allclose(ax[0], (0, 0, 1), atol=1e-12)
    assert tuple(ex["collimation_half_angles"]) == (0.0, 0.0)


def test_bowtie_reader(tmp_path):
    f = read_filters()
    assert DEFAULT_CT_BOWTIE in f
    bt = f[DEFAULT_CT_BOWTIE]
    assert len(bt.angle) == 7 and bt.weight[np.argmin(bt.angle)] == pytest.approx(23.5782)
    bad = tmp_path / "b.json"
    bad.write_text('{"filters": [{"name": "x", "filterdata": [{"angle": "a", "weight": 1}]},'
                   ' {"filterdata": [{"angle": 0.1, "weight": 1}]}, {"name": "ok", "filterdata": [{"angle": 0.0, "weight": 2.0}]}]}')
    got = read_filters(bad)
    assert list(got) == ["ok"]
    assert read_filters(tmp_path / "missing.json") == {}


def test_phantoms():
    ph = phantoms.ct_body((64, 64, 50), (0.6, 0.6, 0.6), seed=5)
    counts = np.bincount(ph.material, minlength=5)
    assert all(counts > 0)
    assert set(np.unique(ph.density)) == {phantoms.AIR_DENSITY, 1.03, 0.95, 0.26, 1.92}
    a = phantoms.ct_body((32, 32, 10), (1.0, 1.0, 1.0), seed=1)
    b = phantoms.ct_body((32, 32, 10), (1.0, 1.0, 1.0), seed=1)
    np.testing.assert_array_equal(a.material, b.material)
    for comp in (phantoms.SOFT_TISSUE, phantoms.ADIPOSE, phantoms.LUNG, phantoms.BONE):
        assert sum(comp.values()) == pytest.approx(1.0, abs=0.002)


def organ_aec(start_deg, stop_deg, ramp_deg, low, compensate):
    f = dx.CTOrganAECFilter()
    f.setUseFilter(True)
    f.setStartAngleDeg(start_deg)
    f.setStopAngleDeg(stop_deg)
    f.setRampAngleDeg(ramp_deg)
    f.setLowWeightFactor(low)
    f.setCompensateOutside(compensate)
    return f


def test_organ_aec_weight_profile():
    """Low over the arc, linear ramps beside it, high elsewhere; with
    compensation the mean over a turn is 1 (numerical integral over a fine
    angle grid), without it the high weight is 1."""
    f = organ_aec(60.0, 180.0, 20.0, 0.4, True)
    hi = f.maxWeight()
    # arc 120 deg, ramp 20 deg: high = (360 - 0.4 * 140) / (360 - 140)
    assert hi == pytest.approx((360 - 0.4 * 140) / 220, rel=1e-12)
    for deg, want in ((60, 0.4), (120, 0.4), (180, 0.4), (190, 0.4 + 0.5 * (hi - 0.4)), (200, hi), (300, hi),
                      (40, hi), (50, 0.4 + 0.5 * (hi - 0.4)), (60 + 360, 0.4), (-200, 0.4)):
        assert f.weight(math.radians(deg)) == pytest.approx(want, rel=1e-12), deg
    a = np.linspace(0, 2 * np.pi, 360 * 64, endpoint=False)
    assert np.mean([f.weight(x) for x in a]) == pytest.approx(1.0, abs=1e-4)
    g = organ_aec(60.0, 180.0, 20.0, 0.4, False)
    assert g.maxWeight() == 1.0 and g.weight(math.radians(300)) == 1.0
    # an arc through 0: from 300 to 40 degrees
    h = organ_aec(300.0, 40.0, 0.0, 0.5, False)
    assert h.weight(math.radians(350)) == 0.5 and h.weight(math.radians(10)) == 0.5 and h.weight(math.radians(100)) == 1.0
    off = dx.CTOrganAECFilter()
    assert off.weight(1.0) == 1.0 and off.maxWeight() == 1.0


def test_organ_aec_applied_to_spiral_exposures():
    b = dx.CTSpiralBeam((0, 0, -2), (0, 0, 2))
    b.setCollimation(4.0)
    b.setStepAngleDeg(5.0)
    plain = b.exposures()
    oa = b.organAECFilter()
    for k, v in dict(setUseFilter=True, setStartAngleDeg=90.0, setStopAngleDeg=180.0, setRampAngleDeg=10.0,
                     setLowWeightFactor=0.3, setCompensateOutside=True).items():
        getattr(oa, k)(v)
    ex = b.exposures()
    ang = b.start_angle + np.arange(len(ex)) * b.step_angle
    want = plain["weight"] * np.array([oa.weight(x) for x in ang])
    np.testing.assert_allclose(ex["weight"], want, rtol=1e-12)
    assert np.min(ex["weight"]) == pytest.approx(0.3) and np.max(ex["weight"]) == pytest.approx(oa.maxWeight())
    # dual source: each tube by its own angle
    d = dx.CTSpiralDualEnergyBeam((0, 0, -2), (0, 0, 2))
    d.setTubeBoffsetAngleDeg(95.0)
    for k, v in dict(setUseFilter=True, setStartAngleDeg=90.0, setStopAngleDeg=180.0, setLowWeightFactor=0.3).items():
        getattr(d.organAECFilter(), k)(v)
    ex = d.exposures()
    ang = d.start_angle + np.arange(len(ex) // 2) * d.step_angle
    wa = np.array([d.organAECFilter().weight(x) for x in ang])
    wb = np.array([d.organAECFilter().weight(x + math.radians(95.0)) for x in ang])
    np.testing.assert_allclose(ex["weight"][0::2], wa, rtol=1e-12)
    np.testing.assert_allclose(ex["weight"][1::2], wb, rtol=1e-12)


def test_aec_from_water_equivalent_diameter():
    """datacontainer.cpp:42-100: A_w = (mean HU / 1000 + 1) x slice area for CT,
    sum(density) x pixel area for density; D_w = 2 sqrt(A_w / pi); weight
    exp(0.2 D_w); the filter spans the image's z extent about the origin."""
    dims, sp = (20, 16, 3), (0.5, 0.5, 2.0)
    nx, ny, nz = dims
    ct = np.full((nz, ny, nx), -1000.0)   # air
    ct[1] = 0.0                             # water slice: A_w = the slice area 10 x 8 cm2
    ct[2, :8, :] = 1000.0                   # half the slice at +1000 HU: mean 0 HU over the slice... plus air
    ct[2, 8:, :] = -1000.0
    f = dx.CTAECFilter.fromWaterEquivalentDiameter(dims, sp, ct, is_ct=True)
    assert f.start == (0.0, 0.0, -3.0) and f.stop == (0.0, 0.0, 3.0)
    area = nx * ny * sp[0] * sp[1]
    np.testing.assert_allclose(f.weights, [1.0, math.exp(0.2 * 2 * math.sqrt(area / math.pi)),
                                           math.exp(0.2 * 2 * math.sqrt(area / math.pi))], rtol=1e-12)
    rho = np.zeros((nz, ny, nx))
    rho[0, :, :4] = 1.0   # a 2 cm x 8 cm block of water
    g = dx.CTAECFilter.fromWaterEquivalentDiameter(dims, sp, rho, is_ct=False)
    np.testing.assert_allclose(g.weights, [math.exp(0.2 * 2 * math.sqrt(16.0 / math.pi)), 1.0, 1.0], rtol=1e-12)
