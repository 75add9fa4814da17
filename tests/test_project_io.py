"""OpenDXMC project files (hdf5wrapper.cpp) through include/dxmc_h5.h and
opendxmc_amd/project.py.  CPU only.

The HDF5 library itself is the checker of the format: the multi-threaded
deflate stream must be inflated by HDF5's own deflate filter (zlib, Adler-32
checked) to the bytes written, and h5dump -- HDF5's own tool, not our code --
must show the layout the reference's saveArray / saveAttribute calls produce
(extent slowest axis first, one chunk of the whole extent with deflate level 6
above 128 elements, contiguous otherwise, variable-length strings, scalar or
1-d attributes).  Project and beam round trips check the reference's paths and
its load quirks.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

import opendxmc_amd as dx
from opendxmc_amd import project as PR

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def h5dump(path, *args) -> str:
    if H5DUMP is None:
        pytest.skip("h5dump is not installed")
    return subprocess.run([H5DUMP, *args, str(path)], capture_output=True, text=True, check=True).stdout


def block(text: str, name: str) -> str:
    """The h5dump block of the object called name."""
    i = text.index(f'"{name}"')
    depth, j = 0, text.index("{", i)
    for k in range(j, len(text)):
        depth += text[k] == "{"
        depth -= text[k] == "}"
        if depth == 0:
            return text[i:k + 1]
    return text[i:]


def test_arrays_strings_attributes_round_trip(tmp_path):
    p = tmp_path / "a.h5"
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((5, 7, 11))
    with PR.H5File(p, "w") as f:
        f.write_array("/g/h/vol", vol, 6)
        f.write_array("u8", np.arange(200, dtype=np.uint8) % 7, 6)
        f.write_array("u64", np.array([3, 4, 5], dtype=np.uint64))
        f.write_array("f32", np.linspace(0, 1, 9, dtype=np.float32))
        f.write_strings("names", ["air", "", "soft tissue", "x" * 300])
        f.write_strings("empty", [])
        f.write_attribute("beams/B/1", "scalar", 2.5)
        f.write_attribute("beams/B/1", "vec", [1.0, 2.0, 3.0])
        f.write_attribute("beams/B/1", "count", np.uint64(12), dtype=np.uint64)
        f.create_group("only/a/group")
        with pytest.raises(dx._native.NativeError):
            f.write_array("u64", np.zeros(2))                   # exists
    with PR.H5File(p, "r") as f:
        assert f.list("/") == sorted(["g", "u8", "u64", "f32", "names", "empty", "beams", "only"])
        assert (f.kind("g/h"), f.kind("/g/h/vol"), f.kind("missing"), f.kind("g/missing/x")) == (1, 2, 0, 0)
        assert np.array_equal(f.read_array("g/h/vol"), vol)
        assert f.shape("g/h/vol") == ((5, 7, 11), 1, 8)
        assert np.array_equal(f.read_array("u8", np.uint8), np.arange(200, dtype=np.uint8) % 7)
        assert np.array_equal(f.read_array("u64", np.float64), [3.0, 4.0, 5.0])   # converted like H5Dread
        assert np.array_equal(f.read_array("f32", np.float64), np.linspace(0, 1, 9, dtype=np.float32))
        assert f.read_strings("names") == ["air", "", "soft tissue", "x" * 300]
        assert f.read_strings("empty") == []
        v, r = f.read_attribute("beams/B/1", "scalar")
        assert (list(v), r) == ([2.5], 0)
        v, r = f.read_attribute("beams/B/1", "vec")
        assert (list(v), r) == ([1.0, 2.0, 3.0], 1)
        assert list(f.read_attribute("beams/B/1", "count", np.uint64)[0]) == [12]
        assert f.read_attribute("beams/B/1", "count") is None            # not a float: loadAttribute's nullopt
        assert f.read_attribute("beams/B/1", "missing") is None
        with pytest.raises(dx._native.NativeError):
            f.write_array("new", np.zeros(3))                               # read-only


@pytest.mark.parametrize("library_filter", [False, True])
def test_multithreaded_deflate_is_read_by_the_hdf5_filter(tmp_path, monkeypatch, library_filter):
    """40 MB over 1 MiB blocks: HDF5's deflate filter inflates the threaded
    stream (dictionary-primed, sync-flushed blocks, combined Adler-32) back to
    the same bytes; the file matches the one HDF5's own filter writes."""
    if library_filter:
        monkeypatch.setenv("DXG_H5_LIBRARY_DEFLATE", "1")
    rng = np.random.default_rng(1)
    a = np.cumsum(rng.integers(-3, 4, 5_000_000)).astype(np.float64) * 0.25
    p = tmp_path / f"big{int(library_filter)}.h5"
    with PR.H5File(p, "w") as f:
        f.write_array("a", a, 6)
        ms, threads = PR.last_deflate()
        assert ms > 0 and (threads == 1 if library_filter else threads >= 1)
    with PR.H5File(p, "r") as f:
        assert np.array_equal(f.read_array("a"), a)
    meta = block(h5dump(p, "-p", "-H"), "a")
    assert "CHUNKED ( 5000000 )" in meta and "COMPRESSION DEFLATE { LEVEL 6 }" in meta


def test_project_layout_matches_the_reference(tmp_path):
    data = PR.synthetic_project((12, 10, 6))
    data.aec = dx.CTAECFilter((0.0, 0.0, -5.0), (0.0, 0.0, 5.0), np.linspace(0.5, 1.5, 20))
    p = tmp_path / "project.h5"
    with PR.HDF5Wrapper(p, PR.HDF5Wrapper.WriteOver) as w:
        assert w.save(data)
    text = h5dump(p, "-p", "-H")
    for name in ("densityarray", "ctarray", "dosearray", "dosevariancearray", "doseeventcountarray"):
        b = block(text, name)
        assert "H5T_IEEE_F64LE" in b and "SIMPLE { ( 6, 10, 12 ) / ( 6, 10, 12 ) }" in b, name
        assert "CHUNKED ( 6, 10, 12 )" in b and "DEFLATE { LEVEL 6 }" in b, name
    for name in ("materialarray", "organarray"):
        assert "H5T_STD_U8LE" in block(text, name)
    assert "H5T_STD_U64LE" in block(text, "dimensions") and "CONTIGUOUS" in block(text, "dimensions")
    assert "CONTIGUOUS" in block(text, "spacing")
    assert "CONTIGUOUS" in block(text, "aecweights")                  # 20 values: not compressed
    assert "STRSIZE H5T_VARIABLE" in block(text, "materialnames")
    values = h5dump(p, "-d", "/dimensions")
    assert "12, 10, 6" in values
    comp = h5dump(p, "-d", "/materialcomposition")
    assert '"H0.100000O0.900000"' in comp


def test_project_round_trip_and_load_quirks(tmp_path):
    data = PR.synthetic_project((16, 12, 8))
    data.aec = dx.CTAECFilter((0.0, 0.0, -5.0), (0.0, 0.0, 5.0), np.linspace(0.5, 1.5, 200))
    p = tmp_path / "project.h5"
    with PR.HDF5Wrapper(p, PR.HDF5Wrapper.WriteOver) as w:
        w.save(data)
    with PR.HDF5Wrapper(p, PR.HDF5Wrapper.ReadOnly) as r:
        back = r.load()
    assert back.dimensions == (16, 12, 8) and back.spacing == (0.1, 0.1, 0.1)
    for f in ("density", "ct", "material", "organ", "dose", "dose_variance", "dose_count"):
        assert np.array_equal(getattr(back, f), getattr(data, f)), f
    assert back.organ_names == data.organ_names
    assert back.materials == [(n, {1: 0.1, 8: 0.9}) for n, _ in data.materials]
    assert back.aec is None                        # saved as "aecweights", loaded from "weights"
    # a project needs dimensions, spacing, materials and density
    q = tmp_path / "bad.h5"
    with PR.HDF5Wrapper(q, PR.HDF5Wrapper.WriteOver) as w:
        d2 = PR.synthetic_project((4, 4, 4))
        d2.density = None
        w.save(d2)
    with PR.HDF5Wrapper(q, PR.HDF5Wrapper.ReadOnly) as r:
        assert r.load() is None
    assert PR.HDF5Wrapper(tmp_path / "missing.h5", PR.HDF5Wrapper.ReadOnly).load() is None


def test_composition_strings():
    water = {1: 0.111894, 8: 0.888106}
    s = PR.composition_string(water)
    assert s == "H0.111894O0.888106"
    assert PR.parse_composition(s) == water
    assert PR.parse_composition("Ca0.2C0.8") == {20: 0.2, 6: 0.8}
    with pytest.raises(ValueError):
        PR.parse_composition("Xx1.0")


def test_beams_round_trip_with_the_reference_quirks(tmp_path):
    dxb = dx.DXBeam({13: 2.0, 29: 0.1})
    dxb.rotation_center = (1.0, 2.0, 3.0)
    dxb.angles = (np.radians(30.0), np.radians(-10.0))
    dxb.setDAPvalue(12.5)
    dxb.setNumberOfExposures(3)
    dxb.setNumberOfParticlesPerExposure(777)
    spiral = dx.CTSpiralBeam((0, 0, -5), (0, 0, 5), {13: 6.0, 50: 0.4})
    spiral.setPitch(1.2)
    spiral.setCTDIvol(8.0)
    spiral.setCTDIdiameter(16.0)
    spiral.setStepAngleDeg(3.0)
    oaec = spiral.organAECFilter()
    oaec.setUseFilter(True)
    oaec.setLowWeightFactor(0.4)
    oaec.setStartAngle(0.5)
    oaec.setStopAngle(1.5)
    oaec.setRampAngle(0.2)
    seq = dx.CTSequentialBeam((1.0, 2.0, 3.0))
    seq.setNumberOfSlices(4)
    seq.setCTDIw(11.0)
    seq.setCTDIdiameter(16.0)
    cbct = dx.CBCTBeam()
    cbct.setStartAngleDeg(-45.0)
    cbct.setDAPvalue(3.0)
    dual = dx.CTSpiralDualEnergyBeam()
    p = tmp_path / "beams.h5"
    with PR.HDF5Wrapper(p, PR.HDF5Wrapper.WriteOver) as w:
        for b in (dxb, spiral, seq, cbct, dual, spiral):
            assert w.saveBeam(b)
    text = h5dump(p, "-H")
    for g in ("DXBeams", "CTSpiralBeams", "CTSequentialBeams", "CBCTBeams", "CTSpiralDualEnergyBeams"):
        assert f'GROUP "{g}"' in text
    assert "DATASPACE  SCALAR" in block(text, "DAPvalue")
    assert "H5T_STD_U64LE" in block(text, "particles_per_exposure")
    with PR.HDF5Wrapper(p, PR.HDF5Wrapper.ReadOnly) as r:
        beams = r.loadBeams()
    kinds = [type(b).__name__ for b in beams]
    assert kinds == ["DXBeam", "CTSpiralBeam", "CTSpiralBeam", "CBCTBeam", "CTSequentialBeam"]   # no dual-energy
    d, s1, s2, c, q = beams
    assert d.rotation_center == (1.0, 2.0, 3.0) and np.allclose(np.degrees(d.angles), (30.0, -10.0))
    assert (d.DAPvalue(), d.n_exposures, d.numberOfParticlesPerExposure) == (12.5, 3, 777)
    assert d.tube.filtration == {13: 2.0, 29: 0.1, 50: 0.0}
    assert np.allclose(d.half_angles, dxb.half_angles)
    for s in (s1, s2):
        assert s.start == (0.0, 0.0, -5.0) and s.pitch == 1.2 and s.CTDIvol() == 8.0
        assert s.CTDIdiameter() == 32.0               # saved "CTDIdiameter", loaded "CTDIDiameter"
        assert np.isclose(np.degrees(s.step_angle), 3.0) and s.tube.filtration[50] == 0.4
        o = s.organAECFilter()
        assert (o.useFilter(), o.compensateOutside(), o.lowWeight()) == (True, False, 0.4)
        assert (o.startAngle(), o.stopAngle(), o.rampAngle()) == (0.5, 1.5, 0.2)
    assert q.position == (0.0, 0.0, 0.0)              # saved "start_position", loaded "position"
    assert (q.slices, q.CTDIw(), q.CTDIdiameter()) == (4, 11.0, 32.0)
    assert np.isclose(c.start_angle, np.radians(-45.0)) and c.DAPvalue() == 3.0
