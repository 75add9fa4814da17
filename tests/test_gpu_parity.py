"""HIP engine vs oracle on the GPU, through the C-ABI (libdxmc_hip.so).

* Deterministic traversal (box entry + voxel lookup) is bit-exact against the
  oracle's FP32 traversal on identical inputs, including rays on voxel faces.
* Transport: per-voxel statistical parity and total-energy parity (tests/parity.py)
  for the pencil-beam water cube, a CT spiral with spectrum, bowtie and AEC
  on a small body phantom, and the DX, CBCT, sequential and dual-source
  families; both correction levels.
* Exact energy bookkeeping on the device counters, cancellation mid-run and
  the variance estimator against the spread of independent seeds.
"""
from __future__ import annotations

import math
import os
import threading

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import phantoms
from opendxmc_amd.bowtie import DEFAULT_CT_BOWTIE, read_filters
from parity import compare, make_world, run_both
from rays import face_rays, random_rays

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,spacing", [((13, 9, 7), (0.25, 0.5, 0.3)), ((64, 64, 64), (0.1, 0.1, 0.1)),
                                          ((256, 256, 200), (0.15, 0.15, 0.15)), ((512, 512, 16), (0.08, 0.08, 0.08))])
def test_traverse_bit_exact(oracle, dims, spacing):
    """The tracking's own entry, advance and voxel lookup (boxEntry, advance,
    voxelIndex, wordMaterial of transport.hip, replayed by traverse_kernel)
    against the oracle's FP32 restatement: random rays, axis rays on faces,
    and lookups at and within 3 ulps of every voxel face of each axis."""
    rng = np.random.default_rng(3)
    n = int(np.prod(dims))
    mat = rng.integers(0, 3, n).astype(np.uint8)
    ph = phantoms.Phantom(dims, spacing, np.ones(n), mat, [phantoms.AIR, phantoms.WATER, phantoms.BONE], ["a", "w", "b"])
    _, grid = make_world(ph, 1)
    sp = np.array(spacing, dtype=np.float32)
    pos, dirs, steps = random_rays(rng, dims, sp, 4096, 24)
    fpos, fdirs, fsteps = face_rays(rng, dims, sp, 24)
    for p, d, st in ((pos, dirs, steps), (fpos, fdirs, fsteps)):
        gi, gm = grid.traverse(p, d, st)
        oi, om = oracle.traverse_f32(dims, sp, mat, p, d, st)
        assert (gi >= 0).sum() > len(p)  # the rays do visit the grid
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gm, om)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_pencil_water_cube(oracle, level):
    ph = phantoms.water_cube(64, 0.1)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
    beam.setNumberOfParticlesPerExposure(1_000_000)
    g, o, hg, ho, gc, oc = run_both(ph, beam, level, oracle=oracle)
    v = compare(g, o, hg, ho)
    print(v, flush=True)
    assert hg == ho == 1_000_000
    assert v.ok, str(v)
    # energy bookkeeping on the device: emitted = deposited + escaped
    assert gc["histories"] == hg
    assert abs(gc["deposited_energy"] + gc["escaped_energy"] - gc["emitted_energy"]) <= 1e-5 * gc["emitted_energy"]
    assert abs(gc["deposited_energy"] - g[0].sum()) <= 1e-6 * g[0].sum()
    # event counts agree statistically with the oracle
    assert abs(gc["real_events"] / hg - oc["real_events"] / ho) <= 0.01 * oc["real_events"] / ho


def small_body():
    return phantoms.ct_body((64, 64, 50), (0.6, 0.6, 0.6), seed=5)


def test_ct_spiral_spectrum_bowtie_aec(oracle):
    ph = small_body()
    beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(15)
    beam.setPitch(1.0)
    beam.setTubeVoltage(120)
    beam.addTubeFiltrationMaterial(13, 6.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setAECFilter(dx.CTAECFilter((0, 0, -12), (0, 0, 12), (1.0, 1.4, 2.0, 1.4, 1.0)))
    n_ex = beam.numberOfExposures()
    beam.setNumberOfParticlesPerExposure(3_000_000 // n_ex)
    g, o, hg, ho, gc, oc = run_both(ph, beam, 1, oracle=oracle)
    v = compare(g, o, hg, ho)
    print(v, flush=True)
    assert v.ok, str(v)
    assert abs(gc["deposited_energy"] + gc["escaped_energy"] - gc["emitted_energy"]) <= 1e-5 * gc["emitted_energy"]


def test_ct_spiral_level2_iodine_insert(oracle):
    """Correction level 2: impulse-approximation Compton and K fluorescence,
    with an iodinated insert (5 % I by mass) and bone so K lines of I and Ca fly."""
    ph = small_body()
    mats = list(ph.materials) + [{53: 0.05, **{z: w * 0.95 for z, w in phantoms.SOFT_TISSUE.items()}}]
    mat = ph.material.reshape(ph.dims[::-1]).copy()
    dens = ph.density.reshape(ph.dims[::-1]).copy()
    zz, yy, xx = np.meshgrid(*[(np.arange(d) + 0.5 - d / 2) * s for d, s in zip(ph.dims[::-1], ph.spacing[::-1])],
                             indexing="ij")
    insert = ((xx - 4.0) ** 2 + yy ** 2 < 2.5 ** 2) & (mat == 1)
    mat[insert] = 5
    dens[insert] = 1.05
    ph = phantoms.Phantom(ph.dims, ph.spacing, dens.ravel(), mat.ravel(), mats, ph.names + ["iodine insert"])
    beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(15)
    beam.setTubeVoltage(120)
    beam.addTubeFiltrationMaterial(13, 6.0)
    beam.setNumberOfParticlesPerExposure(3_000_000 // beam.numberOfExposures())
    g, o, hg, ho, gc, oc = run_both(ph, beam, 2, oracle=oracle)
    v = compare(g, o, hg, ho)
    print(v, flush=True)
    assert v.ok, str(v)
    assert abs(gc["deposited_energy"] + gc["escaped_energy"] - gc["emitted_energy"]) <= 1e-5 * gc["emitted_energy"]


@pytest.mark.parametrize("family", ["dx", "cbct", "sequential", "dual"])
def test_beam_families(oracle, family):
    ph = small_body()
    if family == "dx":
        beam = dx.DXBeam({13: 2.0, 29: 0.1})
        beam.tube.voltage = 80.0
        beam.setCollimation((30.0, 30.0))
        beam.angles = (math.radians(30.0), 0.0)
        beam.setNumberOfParticlesPerExposure(2_000_000)
    elif family == "cbct":
        beam = dx.CBCTBeam()
        beam.setStepAngleDeg(10.0)
        beam.setNumberOfParticlesPerExposure(2_000_000 // beam.numberOfExposures())
    elif family == "sequential":
        beam = dx.CTSequentialBeam((0, 0, -4.0))
        beam.setNumberOfSlices(3)
        beam.setSliceSpacing(4.0)
        beam.setStepAngleDeg(20.0)
        beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
        beam.setNumberOfParticlesPerExposure(2_000_000 // beam.numberOfExposures())
    else:
        beam = dx.CTSpiralDualEnergyBeam((0, 0, -8), (0, 0, 8))
        beam.setCollimation(4.0)
        beam.setStepAngleDeg(20.0)
        beam.setPitch(2.0)
        beam.setTubeAVoltage(100)
        beam.setTubeBVoltage(140)
        beam.addTubeFiltrationMaterial(13, 6.0)
        beam.tubeB.filtration = {13: 6.0, 50: 0.4}
        beam.setRelativeMasTubeB(0.7)
        beam.setBowtieFilterA(read_filters()[DEFAULT_CT_BOWTIE])
        beam.setNumberOfParticlesPerExposure(2_000_000 // beam.numberOfExposures())
    g, o, hg, ho, gc, oc = run_both(ph, beam, 1, oracle=oracle)
    v = compare(g, o, hg, ho)
    print(family, v, flush=True)
    assert v.ok, str(v)


def test_spiral_with_wed_and_organ_aec(oracle):
    """A CT spiral whose tube current follows the water-equivalent diameter
    per slice (datacontainer.cpp:42-100) and an organ AEC arc with ramps and
    compensation (beamsettingsmodel.cpp:349-437): the exposure weights differ
    by a factor of several across the scan, and the transport still matches
    the oracle."""
    ph = small_body()
    hz = 0.5 * ph.dims[2] * ph.spacing[2]
    beam = dx.CTSpiralBeam((0, 0, -hz), (0, 0, hz), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setPitch(1.0)
    beam.setStepAngleDeg(15.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setAECFilter(dx.CTAECFilter.fromWaterEquivalentDiameter(ph.dims, ph.spacing, ph.density, is_ct=False))
    o = beam.organAECFilter()
    o.setUseFilter(True)
    o.setStartAngleDeg(45.0)
    o.setStopAngleDeg(135.0)
    o.setRampAngleDeg(20.0)
    o.setLowWeightFactor(0.3)
    o.setCompensateOutside(True)
    beam.setNumberOfParticlesPerExposure(3_000_000 // beam.numberOfExposures())
    w = beam.exposures()["weight"]
    assert w.max() / w.min() > 3.0
    g, og, hg, ho, gc, oc = run_both(ph, beam, 1, oracle=oracle)
    v = compare(g, og, hg, ho)
    print("wed+organ aec", v, flush=True)
    assert v.ok, str(v)


def test_cancel_mid_run():
    ph = phantoms.water_cube(32, 0.2)
    world, grid = make_world(ph, 1)
    beam = dx.PencilBeam((0.1, 0.1, -10.0), (0, 0, 1), 60.0)
    beam.setNumberOfParticlesPerExposure(400_000_000)
    prog = dx.TransportProgress()
    os.environ["DXG_CHUNK"] = str(1 << 20)
    try:
        def stopper():
            import time
            t0 = time.time()
            while prog.progress()[0] == 0 and time.time() - t0 < 60:
                time.sleep(0.001)
            prog.setStopSimulation()
        th = threading.Thread(target=stopper)
        th.start()
        finished = dx.Transport(seed=3)(world, beam, prog, True)
        th.join()
    finally:
        del os.environ["DXG_CHUNK"]
    assert finished is False
    assert not prog.continueSimulation()
    n, total = prog.progress()
    assert total == 400_000_000 and 0 < n < total


@pytest.mark.parametrize("level", [0, 1])
def test_woodcock_beer_lambert(level):
    """Uncollided fraction through water 2 cm / bone 1 cm / lung 3 cm equals
    exp(-sum mu rho t) on the device (FP32 tables, binned majorant)."""
    dims = (8, 8, 60)
    mat = np.zeros(dims[::-1], dtype=np.uint8)
    dens = np.zeros(dims[::-1])
    mat[0:20], dens[0:20] = 1, 1.0
    mat[20:30], dens[20:30] = 2, 1.92
    mat[30:60], dens[30:60] = 3, 0.26
    comps = [phantoms.AIR, phantoms.WATER, phantoms.BONE, phantoms.LUNG]
    ph = phantoms.Phantom(dims, (0.1, 0.1, 0.1), dens.ravel(), mat.ravel(), comps, ["a", "w", "b", "l"])
    _, grid = make_world(ph, level)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
    n = 4_000_000
    beam.setNumberOfParticlesPerExposure(n)
    world = dx.World()
    world.items.append(grid)
    assert dx.Transport(seed=21)(world, beam)
    c = grid.counters()
    mu = [dx.Material(m).attenuationValues([60.0])[0, 3] for m in comps]
    expected = math.exp(-(mu[1] * 2.0 + mu[2] * 1.92 * 1.0 + mu[3] * 0.26 * 3.0))
    got = c["uncollided"] / n
    se = math.sqrt(expected * (1 - expected) / n)
    print(f"uncollided {got:.6f} expected {expected:.6f} ({(got - expected) / se:+.2f} SE)", flush=True)
    assert abs(got - expected) < 4 * se + 2e-4 * expected   # + FP32 table interpolation allowance


def test_outer_medium_flights_match_woodcock_steps():
    """Outside the box of its non-air voxels the body phantom is uniform air.
    Photons there sample their flights exactly in air (to a real collision,
    the box's face or the grid's edge) instead of taking Woodcock steps with
    the bone-set majorant.  Both must sample the same transport: the two runs
    agree voxel by voxel within their standard errors, in total energy, in the
    energy deposited in the outer air and in the fraction of photons that
    leave without interacting; the exact flights take fewer Woodcock steps."""
    ph = small_body()
    beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(15)
    beam.setPitch(1.0)
    beam.setTubeVoltage(120)
    beam.addTubeFiltrationMaterial(13, 6.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setNumberOfParticlesPerExposure(4_000_000 // beam.numberOfExposures())
    world, grid = make_world(ph, 1)
    outer = (ph.material == ph.material.flat[0]) & (ph.density == ph.density.flat[0])
    assert 0 < outer.sum() < outer.size
    runs = {}
    old = os.environ.get("DXG_OUTER")
    try:
        for mode, env in (("exact", "1"), ("woodcock", "0")):
            os.environ["DXG_OUTER"] = env
            grid.clearScores()
            assert dx.Transport(seed=31 if mode == "exact" else 32)(world, beam, None, False)
            e, e2, _ = grid.energyScored()
            runs[mode] = (e, e2, grid.histories(), grid.counters())
    finally:
        if old is None:
            del os.environ["DXG_OUTER"]
        else:
            os.environ["DXG_OUTER"] = old
    (e1, s1, h1, c1), (e0, s0, h0, c0) = runs["exact"], runs["woodcock"]
    v = compare((e1, s1), (e0, s0), h1, h0)
    print(v, flush=True)
    assert v.ok, str(v)
    print("woodcock steps per history", c1["steps"] / h1, "against", c0["steps"] / h0, flush=True)
    assert c1["steps"] / h1 < 0.9 * c0["steps"] / h0
    for c in (c1, c0):
        assert abs(c["deposited_energy"] + c["escaped_energy"] - c["emitted_energy"]) <= 1e-5 * c["emitted_energy"]
    # energy deposited in the outer air, per history, within 4 standard errors
    a1, a0 = e1[outer].sum() / h1, e0[outer].sum() / h0
    va = (s1[outer].sum() / h1 - a1 * a1) / h1 + (s0[outer].sum() / h0 - a0 * a0) / h0
    print("outer-air energy per history", a1, a0, "z", (a1 - a0) / math.sqrt(va), flush=True)
    assert a1 > 0 and abs(a1 - a0) <= 4.0 * math.sqrt(va)
    # photons leaving without any interaction (binomial standard error)
    u1, u0 = c1["uncollided"] / h1, c0["uncollided"] / h0
    su = math.sqrt(u1 * (1 - u1) / h1 + u0 * (1 - u0) / h0)
    print("uncollided fraction", u1, u0, "z", (u1 - u0) / su, flush=True)
    assert abs(u1 - u0) <= 4.0 * su


@pytest.mark.parametrize("edge", [8, 16])
def test_macro_cell_tracking_matches_oracle(oracle, edge):
    """Macro-cell majorants (forced on a small grid with DXG_CELL at build
    time): the walk through cells of ratio r_c * majorant, tentative collisions
    inside a cell accepted with mu / mu_c, gives the oracle's doses; the walk
    does cross cells."""
    ph = small_body()
    beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(15)
    beam.setPitch(1.0)
    beam.setTubeVoltage(120)
    beam.addTubeFiltrationMaterial(13, 6.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setNumberOfParticlesPerExposure(3_000_000 // beam.numberOfExposures())
    old = os.environ.get("DXG_CELL")
    os.environ["DXG_CELL"] = str(edge)
    try:
        g, o, hg, ho, gc, oc = run_both(ph, beam, 1, oracle=oracle)
    finally:
        if old is None:
            del os.environ["DXG_CELL"]
        else:
            os.environ["DXG_CELL"] = old
    v = compare(g, o, hg, ho)
    print(v, "crossings per history", gc["cell_crossings"] / hg, flush=True)
    assert v.ok, str(v)
    assert gc["cell_crossings"] > 0
    assert abs(gc["deposited_energy"] + gc["escaped_energy"] - gc["emitted_energy"]) <= 1e-5 * gc["emitted_energy"]


def test_results_do_not_depend_on_the_phase_schedule():
    """The per-history random streams are consumed in the same order whatever
    the phase schedule, so runs of one seed with very different refill,
    interaction and outer-flight thresholds (read at each run) give identical
    event counters and tallies equal up to the order of the FP64 atomics."""
    ph = small_body()
    beam = dx.CTSpiralBeam((0, 0, -6), (0, 0, 6))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(20)
    beam.setTubeVoltage(120)
    beam.addTubeFiltrationMaterial(13, 6.0)
    beam.setBowtieFilter(read_filters()[DEFAULT_CT_BOWTIE])
    beam.setNumberOfParticlesPerExposure(1_000_000 // beam.numberOfExposures())
    world, grid = make_world(ph, 1)
    keys = ("DXG_REFILL", "DXG_INTERACT", "DXG_OUTERT")
    old = {k: os.environ.get(k) for k in keys}
    results = []
    try:
        for refill, interact, outer in ((16, 24, 16), (1, 1, 1), (64, 64, 64), (5, 40, 3)):
            os.environ.update(DXG_REFILL=str(refill), DXG_INTERACT=str(interact), DXG_OUTERT=str(outer))
            grid.clearScores()
            assert dx.Transport(seed=123)(world, beam, None, False)
            e, e2, n = grid.energyScored()
            results.append((grid.counters(), e, e2, n))
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    c0, e0, s0, n0 = results[0]
    for c, e, s2, n in results[1:]:
        for k in ("histories", "steps", "real_events", "deposits", "uncollided"):
            assert c[k] == c0[k], k
        np.testing.assert_array_equal(n, n0)
        np.testing.assert_allclose(e, e0, rtol=1e-9, atol=1e-12 * e0.max())
        np.testing.assert_allclose(s2, s0, rtol=1e-9, atol=1e-12 * s0.max())


@pytest.mark.parametrize("small_grid", [False, True])
def test_extreme_launch_settings_leave_events_unchanged(small_grid):
    """The source queue and the phase thresholds decide only which lane runs a
    history and when, never what it draws: refill at every dead lane or only
    when all 64 are dead, interaction and outer flights at 1 or 64 parked
    lanes, pool claims of 64 or 4096 histories, and (small grid) the tally
    merge on and off all give the same event counts; the energy scores agree
    up to the order of FP64 additions."""
    if small_grid:
        ph = phantoms.water_cube(16, 0.25)
        beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
        beam.setNumberOfParticlesPerExposure(3_000_000)
    else:
        ph = small_body()
        beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12), {13: 6.0})
        beam.setCollimation(4.0)
        beam.setStepAngleDeg(10.0)
        beam.setNumberOfParticlesPerExposure(4_000_000 // beam.numberOfExposures())
    settings = [
        {},
        {"DXG_REFILL": "1", "DXG_INTERACT": "1", "DXG_OUTERT": "1", "DXG_BATCH": "64"},
        {"DXG_REFILL": "64", "DXG_INTERACT": "64", "DXG_OUTERT": "64", "DXG_BATCH": "4096"},
        {"DXG_REFILL": "8", "DXG_INTERACT": "40", "DXG_OUTERT": "4", "DXG_TALLY_MERGE": "0" if small_grid else "1"},
    ]
    ref = None
    for env in settings:
        world, grid = make_world(ph, 1)
        os.environ.update(env)
        try:
            assert dx.Transport(seed=23)(world, beam, None, False)
        finally:
            for k in env:
                os.environ.pop(k, None)
        c = grid.counters()
        e, e2, n = grid.energyScored()
        print(f"[settings] {env or 'defaults'}: refill {c['refill_threshold']}, interact {c['interact_threshold']}, "
              f"outer {c['outer_threshold']}, batch {c['batch']}, steps {c['steps']}, deposits {c['deposits']}", flush=True)
        got = ({k: c[k] for k in ("histories", "steps", "real_events", "deposits", "uncollided")}, e, e2, n)
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0], (env, got[0], ref[0])
        np.testing.assert_array_equal(got[3], ref[3])
        np.testing.assert_allclose(got[1], ref[1], rtol=1e-9, atol=1e-12 * ref[1].max())
        np.testing.assert_allclose(got[2], ref[2], rtol=1e-9, atol=1e-12 * ref[2].max())


def test_autotuned_launch_settings_leave_results_unchanged():
    """The first run of a beam layout on a world build times each candidate
    launch setting on short launches and keeps the fastest; the events of the
    run are those of a run with the default settings (DXG_AUTOTUNE=0), and the
    next run reuses the kept settings without timing again."""
    ph = small_body()
    beam = dx.CTSpiralBeam((0, 0, -12), (0, 0, 12), {13: 6.0})
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(10.0)
    beam.setNumberOfParticlesPerExposure(120_000_000 // beam.numberOfExposures())
    counts = {}
    for mode in ("tuned", "default"):
        world, grid = make_world(ph, 1)
        if mode == "default":
            os.environ["DXG_AUTOTUNE"] = "0"
        try:
            assert dx.Transport(seed=5)(world, beam, None, False)
            c = grid.counters()
            print(f"[autotune] {mode}: launches {c['kernel_launches']}, refill {c['refill_threshold']}, interact "
                  f"{c['interact_threshold']}, outer {c['outer_threshold']}, batch {c['batch']}, tuned {c['autotuned']}",
                  flush=True)
            counts[mode] = {k: c[k] for k in ("histories", "steps", "real_events", "deposits", "uncollided")}
            if mode == "tuned":
                # five candidate settings, each timed twice, then the rest of the run
                assert c["autotuned"] == 1 and c["kernel_launches"] > 10
                first = (c["refill_threshold"], c["interact_threshold"], c["outer_threshold"], c["batch"])
                assert dx.Transport(seed=6)(world, beam, None, False)
                c2 = grid.counters()
                assert c2["autotuned"] == 1 and c2["kernel_launches"] < c["kernel_launches"]
                assert (c2["refill_threshold"], c2["interact_threshold"], c2["outer_threshold"], c2["batch"]) == first
            else:
                assert c["autotuned"] == 0 and (c["refill_threshold"], c["interact_threshold"]) == (16, 24)
        finally:
            os.environ.pop("DXG_AUTOTUNE", None)
    assert counts["tuned"] == counts["default"]


def test_air_fill_medium_effect_on_the_pencil_cube():
    """The world here ends at the grid (vacuum beyond); DXMClib's World fills
    the space around its items with air (its padding is not in the reference
    tree).  Its effect on the 64^3 water pencil-beam case: the same cube inside
    10 cm of air on every side (air voxels, 0.001205 g/cm3; the source inside
    the air at z = -10 cm) against the bare cube: every voxel above 10 % of the
    maximum agrees within 3 SE and the total dose within 0.5 %."""
    n, pad, h = 64, 100, 0.1
    N = n + 2 * pad
    dens = np.full((N, N, N), 0.001205)
    mat = np.zeros((N, N, N), np.uint8)
    sl = slice(pad, pad + n)
    dens[sl, sl, sl] = 1.0
    mat[sl, sl, sl] = 1
    padded = phantoms.Phantom((N, N, N), (h, h, h), dens.ravel(), mat.ravel(), [phantoms.AIR, phantoms.WATER],
                              ["air", "water"])
    bare = phantoms.water_cube(n, h)
    beam = dx.PencilBeam((0.05, 0.05, -10.0), (0, 0, 1), 60.0)
    beam.setNumberOfParticlesPerExposure(20_000_000)
    res = {}
    for name, ph in (("bare", bare), ("air", padded)):
        world, grid = make_world(ph, 1)
        assert dx.Transport(seed=71)(world, beam, None, False)
        e, e2, _ = grid.energyScored()
        if name == "air":
            e = e.reshape(N, N, N)[sl, sl, sl].ravel()
            e2 = e2.reshape(N, N, N)[sl, sl, sl].ravel()
        res[name] = (e, e2, grid.histories())
    v = compare(res["air"][:2], res["bare"][:2], res["air"][2], res["bare"][2])
    print(f"[fill medium] cube in 10 cm of air vs bare cube: {v}", flush=True)
    assert v.ok, str(v)
