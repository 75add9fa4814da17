"""CT segmentation (ctsegmentationpipeline.cpp:61-166) and the organ dose
table (dosetablepipeline.cpp:34-87) on the device, include/dxmc_pipelines.h.

CPU: the oracle's restatements are pinned by closed forms (water at 0 HU, air
at -1000 HU, a voxel at a material's own CT number gets that material's
density for a single-energy spectrum, the CT number is linear in the spectrum
weights) and by brute force on small inputs; the library's tables match the
oracle's; bad arguments fail before any device work.  GPU: the per-voxel map
is bit-identical to the oracle's on the same tables (special values, ragged
sizes, unaligned device buffers), the table matches the oracle to 1e-12
relative with exact counts and repeats bit for bit, and size-independent
properties hold at 1.3e8 voxels.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from opendxmc_amd import _native as N
from opendxmc_amd import pipelines as P

WATER = P.NIST_MATERIALS["Water, Liquid"]
AIR = P.NIST_MATERIALS["Air, Dry (near sea level)"]
SEG_MATS = [P.NIST_MATERIALS[n][1] for n in P.SEGMENTATION_MATERIALS]
SEG_DENS = [P.SEGMENTATION_DENSITY_OVERRIDE.get(n, P.NIST_MATERIALS[n][0]) for n in P.SEGMENTATION_MATERIALS]
TABLE_RTOL = 1e-12   # sums over voxels in another order


def oracle_tables(oracle, tube):
    e, w = tube.getSpecter()
    return oracle.mean_ct_numbers(e, w, SEG_MATS, SEG_DENS, WATER[1], WATER[0], AIR[1], AIR[0])


def table_arrays(t: N.CTSegmentation):
    n = t.n_materials
    return np.array(t.separators[: n - 1]), np.array(t.attenuation[:n])


def special_hu(t: N.CTSegmentation, n: int, seed: int) -> np.ndarray:
    """Random CT numbers with the separators themselves, their neighbours,
    integers, infinities, NaN and signed zeros mixed in."""
    rng = np.random.default_rng(seed)
    hu = rng.uniform(-1100.0, 2500.0, n)
    sep, _ = table_arrays(t)
    extra = np.concatenate([sep, np.nextafter(sep, -np.inf), np.nextafter(sep, np.inf),
                            [np.nan, np.inf, -np.inf, 0.0, -0.0, -1000.0, -1024.0, 3071.0, 1e300, -1e300]])
    if n > 2 * len(extra):
        hu[rng.choice(n, n // 3, replace=False)] = np.round(hu[:n // 3])
    k = min(n, len(extra))
    hu[rng.choice(n, k, replace=False)] = extra[:k]
    return hu


def phantom_organs(shape, seed=3):
    """Nested ellipsoids (long runs), a speckled region (short runs) and every
    organ index at least once."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*[np.linspace(-1, 1, s) for s in shape], indexing="ij")
    organ = np.zeros(shape, dtype=np.uint8)
    for k, r in enumerate((0.95, 0.8, 0.6, 0.4, 0.2)):
        organ[(x / r) ** 2 + (y / (0.8 * r)) ** 2 + (z / r) ** 2 <= 1.0] = 10 + 17 * k
    speck = (x > 0.3) & (y > 0.3)
    organ[speck] = rng.integers(0, 256, int(speck.sum()), dtype=np.uint8)
    flat = organ.ravel()
    k = min(256, flat.size)
    flat[:k] = np.arange(k, dtype=np.uint8)
    dose = rng.gamma(2.0, 0.5, flat.size) * (1 + flat % 7)
    density = rng.uniform(0.0, 2.0, flat.size)
    density[rng.choice(flat.size, flat.size // 50, replace=False)] = 0.0
    return flat, dose, density


# ------------------------------------------------------------------ CPU

def test_oracle_single_energy_closed_forms(oracle):
    """One energy: water maps to 0 HU and air to -1000 HU exactly, and a voxel
    at a material's own CT number gets back that material's density."""
    for e in (30.0, 60.0, 110.0):
        mats = [AIR[1], WATER[1]] + SEG_MATS[1:]
        dens = [AIR[0], WATER[0]] + SEG_DENS[1:]
        o = oracle.mean_ct_numbers([e], [1.0], mats, dens, WATER[1], WATER[0], AIR[1], AIR[0])
        assert o["hu"][0] == -1000.0
        assert o["hu"][1] == 0.0
        for m, rho in enumerate(dens):
            _, d = oracle.ct_segment([], [o["attenuation"][m]], o["slope"], o["offset"], [o["hu"][m]])
            assert d[0] == pytest.approx(rho, rel=1e-12)


def test_oracle_ct_number_is_linear_in_the_spectrum(oracle):
    """The CT number averages the per-energy ratio (ctsegmentationpipeline.cpp:
    85-90), so it is linear in the spectrum weights -- a ratio of averaged
    attenuations would not be."""
    e = np.array([40.0, 90.0])
    w = np.array([0.3, 0.7])
    args = (SEG_MATS, SEG_DENS, WATER[1], WATER[0], AIR[1], AIR[0])
    two = oracle.mean_ct_numbers(e, w, *args)
    one = [oracle.mean_ct_numbers([x], [1.0], *args) for x in e]
    np.testing.assert_allclose(two["hu"], w[0] * one[0]["hu"] + w[1] * one[1]["hu"], rtol=1e-12)
    np.testing.assert_allclose(two["attenuation"], w[0] * one[0]["attenuation"] + w[1] * one[1]["attenuation"], rtol=1e-12)
    ratio_of_means = 1000 * (two["attenuation"] * SEG_DENS - two["water_attenuation"] * WATER[0]) / (
        two["water_attenuation"] * WATER[0] - two["air_attenuation"] * AIR[0])
    assert np.max(np.abs(ratio_of_means[1:] - two["hu"][1:])) > 1.0


def test_oracle_map_brute_force(oracle):
    """First separator above the CT number wins (also for unsorted
    separators); std::max keeps NaN and -0.0."""
    sep = [5.0, 0.0, 10.0]
    att = [1.0, 2.0, 4.0, 8.0]
    hu = [3.0, -1.0, 5.0, 7.0, 10.0, 11.0, float("nan"), -float("inf"), float("inf")]
    m, d = oracle.ct_segment(sep, att, 2.0, 1.0, hu)
    expect_m = []
    for h in hu:
        i = next((k for k, s in enumerate(sep) if h < s), len(sep))
        expect_m.append(i)
    assert m.tolist() == expect_m == [0, 0, 2, 2, 3, 3, 3, 0, 3]
    for h, mi, di in zip(hu, m, d):
        v = (2.0 * h / 1000.0 + 1.0) / att[mi]
        if math.isnan(v):
            assert math.isnan(di)
        else:
            assert di == (0.0 if v < 0.0 else v)
    _, dz = oracle.ct_segment([], [1.0], 1.0, -0.0, [-0.0])
    assert math.copysign(1.0, dz[0]) == -1.0


def test_oracle_table_brute_force(oracle):
    rng = np.random.default_rng(1)
    organ = rng.integers(0, 6, 500, dtype=np.uint8)
    organ[organ == 4] = 200
    dose = rng.uniform(0, 3, 500)
    dens = rng.uniform(0, 2, 500)
    sp = (0.1, 0.2, 0.3)
    vv = 0.1 * 0.2 * 0.3
    got = oracle.organ_dose_table(organ, dose, dens, sp)
    for i in range(256):
        sel = [k for k in range(500) if organ[k] == i]
        if not sel:
            assert got[i]["n_voxels"] == 0 and got[i]["mass_g"] == 0.0
            continue
        mass = vv * math.fsum(dens[k] for k in sel)
        energy = math.fsum(dose[k] * (vv * dens[k]) for k in sel)
        assert got[i]["n_voxels"] == len(sel)
        assert got[i]["volume_cm3"] == pytest.approx(len(sel) * vv, rel=1e-15)
        assert got[i]["mass_g"] == pytest.approx(mass, rel=1e-14)
        assert got[i]["dose"] == pytest.approx(energy / mass, rel=1e-14)
    # a uniform dose is every organ's dose
    flat = oracle.organ_dose_table(organ, np.full(500, 2.5), dens, sp)
    np.testing.assert_allclose(flat["dose"][flat["n_voxels"] > 0], 2.5, rtol=1e-14)


def test_library_tables_match_oracle(oracle):
    for kv, al, sn in ((120.0, 0.9, 0.0), (80.0, 2.5, 0.0), (140.0, 6.0, 0.4)):
        p = P.CTSegmentationPipeline()
        p.setAqusitionVoltage(kv)
        p.setAlFiltration(al)
        p.setSnFiltration(sn)
        t = p.tables()
        o = oracle_tables(oracle, p.tube())
        n = t.n_materials
        assert n == 5
        np.testing.assert_allclose(t.hu[:n], o["hu"], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(t.separators[: n - 1], o["separators"], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(t.attenuation[:n], o["attenuation"], rtol=1e-12)
        assert t.slope == pytest.approx(o["slope"], rel=1e-12)
        assert t.offset == pytest.approx(o["offset"], rel=1e-12)
        assert t.hu[0] == pytest.approx(-1000.0, abs=1e-9)   # dry air is the scale's air
        assert list(t.hu[:n]) == sorted(t.hu[:n])


def test_segmentation_settings_follow_the_reference():
    p = P.CTSegmentationPipeline()
    assert (p.kv, p.al_mm, p.sn_mm) == (120.0, 0.9, 0.0)
    p.setAlFiltration(-3.0)
    p.setSnFiltration(-1.0)
    assert (p.al_mm, p.sn_mm) == (0.0, 0.0)
    # harder beam: bone's mean attenuation falls faster than water's
    soft, hard = P.CTSegmentationPipeline(), P.CTSegmentationPipeline()
    hard.setAlFiltration(6.0)
    assert hard.tables().hu[4] < soft.tables().hu[4]


def test_pipeline_arguments_fail_before_device_work():
    lib = N.lib()
    t = P.CTSegmentationPipeline().tables()
    bad = N.CTSegmentation()
    hu = np.zeros(4)
    mat = np.zeros(4, dtype=np.uint8)
    dens = np.zeros(4)
    with pytest.raises(N.NativeError) as e:
        N.check(lib.dxg_ct_segment(C.byref(bad), N.ptr(hu, C.c_double), 4, N.ptr(mat, C.c_uint8), N.ptr(dens, C.c_double)))
    assert e.value.code == N.DXG_ERR_ARGUMENT
    N.check(lib.dxg_ct_segment(C.byref(t), None, 0, None, None))   # nothing to do
    rows = (N.OrganDose * 256)()
    sp = np.array([0.1, 0.0, 0.1])
    with pytest.raises(N.NativeError) as e:
        N.check(lib.dxg_organ_dose_table(N.ptr(mat, C.c_uint8), N.ptr(hu, C.c_double), N.ptr(dens, C.c_double), 4,
                                         N.ptr(sp, C.c_double), rows))
    assert e.value.code == N.DXG_ERR_ARGUMENT
    sp[1] = 0.1
    table = P.organ_dose_table(np.zeros(0, np.uint8), np.zeros(0), np.zeros(0), sp)
    assert (table["n_voxels"] == 0).all()
    with pytest.raises(ValueError):
        P.organ_dose_table(np.zeros(3, np.uint8), np.zeros(4), np.zeros(4), sp)
    k = P._Keep()
    m = P.Material(WATER[1])._struct(k)
    tube = P.Tube(120.0)._struct(k)
    out = N.CTSegmentation()
    d = np.array([1.0])
    with pytest.raises(N.NativeError) as e:
        N.check(lib.dxg_ct_segmentation_tables(C.byref(tube), C.byref(m), N.ptr(d, C.c_double), 0, C.byref(m), 1.0,
                                               C.byref(m), 1e-3, C.byref(out)))
    assert e.value.code == N.DXG_ERR_ARGUMENT


def test_dose_table_rows_follow_the_reference(monkeypatch):
    """Rows only for named organs with voxels, in organ order, with the
    reference's header."""
    table = np.zeros(256, dtype=N.ORGAN_DOSE_DTYPE)
    table[1] = (10, 1.0, 1.1, 2.0)
    table[3] = (5, 0.5, 0.4, 7.0)
    table[300 % 256] = (1, 0.1, 0.1, 1.0)
    monkeypatch.setattr(P, "organ_dose_table", lambda *a: table)
    rows = P.DoseTablePipeline().updateImageData(np.zeros(1), ["air", "liver", "lung", "bone"], np.zeros(1), np.zeros(1),
                                                 (1, 1, 1))
    assert rows == [("liver", 10, 1.0, 1.1, 2.0), ("bone", 5, 0.5, 0.4, 7.0)]
    assert P.DoseTablePipeline.header("mGy") == ["Name", "# Voxels", "Volume cm3", "Mass g", "Dose mGy"]


# ------------------------------------------------------------------ GPU

def assert_same_bits(a: np.ndarray, b: np.ndarray):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 3, 255, 1_000_003])
def test_segment_matches_oracle_bit_for_bit(oracle, n):
    t = P.CTSegmentationPipeline().tables()
    hu = special_hu(t, n, seed=n)
    sep, att = table_arrays(t)
    m_ref, d_ref = oracle.ct_segment(sep, att, t.slope, t.offset, hu)
    m, d = P.segment(t, hu)
    np.testing.assert_array_equal(m, m_ref)
    assert_same_bits(d, d_ref)
    assert P.last_ms(0) > 0.0


@pytest.mark.gpu
def test_segment_device_buffers_any_alignment(oracle):
    """Device buffers, 16-B aligned (paired loads) and one element off (the
    per-voxel kernel), inputs and outputs."""
    from opendxmc_amd.devmem import DeviceBuffer
    t = P.CTSegmentationPipeline().tables()
    n = 300_001
    hu = special_hu(t, n + 1, seed=7)
    sep, att = table_arrays(t)
    d_hu = DeviceBuffer.from_array(hu)
    d_mat = DeviceBuffer(n + 2)
    d_dens = DeviceBuffer(8 * (n + 1))
    for off_in, off_out in ((0, 0), (1, 0), (0, 1), (1, 1)):
        P.segment_device(t, d_hu.ptr + 8 * off_in, n, d_mat.ptr + off_out, d_dens.ptr + 8 * off_out)
        m_ref, d_ref = oracle.ct_segment(sep, att, t.slope, t.offset, hu[off_in:off_in + n])
        np.testing.assert_array_equal(d_mat.download(np.uint8, n, off_out), m_ref)
        assert_same_bits(d_dens.download(np.float64, n, 8 * off_out), d_ref)


@pytest.mark.gpu
def test_device_entry_points_check_their_buffers():
    """A device buffer shorter than the voxel count is refused before any
    launch (2 MiB buffers: no allocation rounding hides the overrun)."""
    from opendxmc_amd.devmem import DeviceBuffer
    t = P.CTSegmentationPipeline().tables()
    n = 1 << 18
    d_a = DeviceBuffer(8 * n)
    d_b = DeviceBuffer(8 * n)
    with pytest.raises(N.NativeError) as e:
        P.segment_device(t, d_a.ptr, n + 1, d_b.ptr, d_b.ptr)
    assert e.value.code == N.DXG_ERR_ARGUMENT and "smaller" in str(e.value)
    with pytest.raises(N.NativeError) as e:
        P.organ_dose_table_device(d_a.ptr, d_a.ptr, d_b.ptr + 8, n, (1, 1, 1))
    assert e.value.code == N.DXG_ERR_ARGUMENT and "smaller" in str(e.value)


@pytest.mark.gpu
def test_segment_large_sampled_against_oracle(oracle):
    """1.3e8 voxels; the oracle checks a sample of them, and within each
    material the density grows with the CT number."""
    t = P.CTSegmentationPipeline().tables()
    n = (1 << 27) + 3
    rng = np.random.default_rng(5)
    hu = rng.uniform(-1100.0, 2500.0, n)
    m, d = P.segment(t, hu)
    idx = rng.integers(0, n, 200_000)
    idx[-1] = n - 1
    sep, att = table_arrays(t)
    m_ref, d_ref = oracle.ct_segment(sep, att, t.slope, t.offset, hu[idx])
    np.testing.assert_array_equal(m[idx], m_ref)
    assert_same_bits(d[idx], d_ref)
    order = np.argsort(hu[:1_000_000])
    ms, ds = m[:1_000_000][order], d[:1_000_000][order]
    assert (ms[1:] >= ms[:-1]).all()
    same = ms[1:] == ms[:-1]
    assert (ds[1:][same] >= ds[:-1][same]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 63), (1, 1, 65), (3, 7, 11), (40, 96, 128), (61, 97, 131)])
def test_dose_table_matches_oracle(oracle, shape):
    organ, dose, density = phantom_organs(shape)
    sp = (0.11, 0.07, 0.3)
    ref = oracle.organ_dose_table(organ, dose, density, sp)
    got = P.organ_dose_table(organ, dose, density, sp)
    np.testing.assert_array_equal(got["n_voxels"], ref["n_voxels"])
    for f in ("volume_cm3", "mass_g"):
        np.testing.assert_allclose(got[f], ref[f], rtol=TABLE_RTOL, atol=0)
    has_mass = ref["mass_g"] > 0
    np.testing.assert_allclose(got["dose"][has_mass], ref["dose"][has_mass], rtol=TABLE_RTOL, atol=0)
    # organs whose voxels all have zero density: 0/0 as in the reference
    empty_mass = (ref["n_voxels"] > 0) & ~has_mass
    assert np.isnan(got["dose"][empty_mass]).all()
    again = P.organ_dose_table(organ, dose, density, sp)
    assert bytes(again.tobytes()) == bytes(got.tobytes())
    assert P.last_ms(1) > 0.0


@pytest.mark.gpu
def test_dose_table_device_buffers_and_properties():
    """Device buffers give the host path's bits; at 1.3e8 voxels the counts
    add to n, the masses to voxel volume x total density, and a uniform dose is
    every organ's dose."""
    from opendxmc_amd.devmem import DeviceBuffer
    organ, dose, density = phantom_organs((40, 96, 128), seed=9)
    sp = (0.2, 0.2, 0.25)
    host = P.organ_dose_table(organ, dose, density, sp)
    bufs = [DeviceBuffer.from_array(a) for a in (organ, dose, density)]
    dev = P.organ_dose_table_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, organ.size, sp)
    assert dev.tobytes() == host.tobytes()
    del bufs

    n = (1 << 27) + 12345
    rng = np.random.default_rng(11)
    runs = rng.integers(1, 5000, n // 2000 + 2)
    organ_big = np.repeat((np.arange(runs.size) % 256).astype(np.uint8), runs)[:n]
    assert organ_big.size == n
    dens_big = rng.uniform(0.0, 2.0, n)
    table = P.organ_dose_table(organ_big, np.full(n, 3.25), dens_big, sp)
    vv = 0.2 * 0.2 * 0.25
    assert int(table["n_voxels"].sum()) == n
    np.testing.assert_array_equal(table["n_voxels"], np.bincount(organ_big, minlength=256))
    assert float(table["mass_g"].sum()) == pytest.approx(vv * float(np.sum(dens_big)), rel=1e-12)
    present = table["n_voxels"] > 0
    np.testing.assert_allclose(table["dose"][present], 3.25, rtol=1e-12)
