"""OpenDXMC project files: HDF5Wrapper (hdf5wrapper.cpp) over the HDF5 layer
of libdxmc_hip.so (include/dxmc_h5.h).

The layout is the reference's, quirks included, so that files pass between
the two programs:
- volumes are stored slowest axis first ({nz, ny, nx}, hdf5wrapper.cpp:119-129),
  as one deflate-6 chunk when they hold more than 128 elements (:143-150);
- beams live in /beams/<Kind>Beams/<n>, n = 1, 2, ... (:461-469), their
  settings as double / uint64 attributes, scalars as scalar dataspaces;
- load() reads the AEC weights from "weights" although save() writes
  "aecweights" (:449 vs :1143), CT beams read "CTDIDiameter" although they
  write "CTDIdiameter" (:569 vs :869), sequential beams read "position"
  although they write "start_position" (:526 vs :779), and loadBeams() looks
  for "CTDualEnergySpiralBeams" while dual-energy spirals are saved under
  "CTSpiralDualEnergyBeams" (:622 vs :1160): those values do not come back.
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .engine import CBCTBeam, CTSequentialBeam, CTSpiralBeam, CTSpiralDualEnergyBeam, CTAECFilter, DXBeam

_TYPES = {np.dtype(np.float64): 0, np.dtype(np.uint64): 1, np.dtype(np.uint8): 2, np.dtype(np.float32): 3,
          np.dtype(np.int32): 4, np.dtype(np.int64): 5}

# AtomHandler::toSymbol for the elements OpenDXMC materials use (Z = 1..100)
SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y "
           "Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re "
           "Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm").split()


def composition_string(comp: dict[int, float]) -> str:
    """Symbol followed by std::to_string(fraction) per element, in Z order
    (hdf5wrapper.cpp:423-431)."""
    return "".join(f"{SYMBOLS[z - 1]}{frac:f}" for z, frac in sorted(comp.items()))


def parse_composition(text: str) -> dict[int, float]:
    """The inverse of composition_string (Material<5>::parseCompoundStr on
    such strings): element symbols, each followed by its number."""
    out: dict[int, float] = {}
    for sym, num in re.findall(r"([A-Z][a-z]?)([0-9]*\.?[0-9]*(?:[eE][-+]?[0-9]+)?)", text):
        if sym not in SYMBOLS:
            raise ValueError(f"unknown element symbol {sym!r} in {text!r}")
        out[SYMBOLS.index(sym) + 1] = out.get(SYMBOLS.index(sym) + 1, 0.0) + (float(num) if num else 1.0)
    return out


class H5File:
    """One HDF5 file through dxg_h5_*: mode "w" truncates, "r" reads."""

    def __init__(self, path: str, mode: str = "r"):
        h = C.c_void_p()
        fn = N.lib().dxg_h5_create if mode == "w" else N.lib().dxg_h5_open
        N.check(fn(str(path).encode(), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            N.check(N.lib().dxg_h5_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def kind(self, path: str) -> int:
        """0 nothing, 1 group, 2 dataset."""
        return int(N.lib().dxg_h5_kind(self._h, path.encode()))

    def create_group(self, path: str) -> None:
        N.check(N.lib().dxg_h5_create_group(self._h, path.encode()))

    def write_array(self, path: str, a: np.ndarray, deflate: int = -1) -> None:
        a = np.ascontiguousarray(a)
        if a.dtype not in _TYPES:
            raise TypeError(f"unsupported element type {a.dtype}")
        dims = np.array(a.shape or (1,), dtype=np.uint64)
        N.check(N.lib().dxg_h5_write_array(self._h, path.encode(), _TYPES[a.dtype], len(dims),
                                           N.ptr(dims, N.u64), a.ctypes.data_as(C.c_void_p), int(deflate)))

    def write_strings(self, path: str, strings: list[str]) -> None:
        enc = [s.encode() for s in strings]
        arr = (C.c_char_p * max(1, len(enc)))(*enc)
        N.check(N.lib().dxg_h5_write_strings(self._h, path.encode(), arr, len(enc)))

    def write_attribute(self, path: str, name: str, values, dtype=np.float64, scalar: bool | None = None) -> None:
        v = np.ascontiguousarray(np.atleast_1d(values), dtype=dtype)
        if scalar is None:
            scalar = np.ndim(values) == 0
        N.check(N.lib().dxg_h5_write_attribute(self._h, path.encode(), name.encode(), _TYPES[v.dtype],
                                               v.ctypes.data_as(C.c_void_p), v.size, int(bool(scalar))))

    def shape(self, path: str) -> tuple[tuple[int, ...], int, int]:
        """(extent, HDF5 type class, element size; 0 for variable-length strings)."""
        rank, cls, size = C.c_int(), C.c_int(), N.u64()
        dims = np.zeros(32, dtype=np.uint64)
        N.check(N.lib().dxg_h5_dataset_shape(self._h, path.encode(), C.byref(rank), N.ptr(dims, N.u64), C.byref(cls),
                                             C.byref(size)))
        return tuple(int(d) for d in dims[: rank.value]), cls.value, int(size.value)

    def read_array(self, path: str, dtype=np.float64) -> np.ndarray:
        dims, _, _ = self.shape(path)
        out = np.empty(dims, dtype=dtype)
        N.check(N.lib().dxg_h5_read_array(self._h, path.encode(), _TYPES[out.dtype], out.ctypes.data_as(C.c_void_p),
                                          out.size))
        return out

    def _strings(self, fn, path: str) -> list[str]:
        count, nbytes = N.u64(), N.u64()
        N.check(fn(self._h, path.encode(), None, 0, C.byref(count), C.byref(nbytes)))
        buf = C.create_string_buffer(max(1, nbytes.value))
        N.check(fn(self._h, path.encode(), buf, nbytes.value, C.byref(count), C.byref(nbytes)))
        if count.value == 0:
            return []
        return [s.decode() for s in buf.raw[: nbytes.value].split(b"\0")[: count.value]]

    def read_strings(self, path: str) -> list[str]:
        return self._strings(N.lib().dxg_h5_read_strings, path)

    def list(self, path: str = "/") -> list[str]:
        return self._strings(N.lib().dxg_h5_list, path)

    def read_attribute(self, path: str, name: str, dtype=np.float64):
        """(values, rank) or None when absent or not an 8-byte float / integer."""
        t = _TYPES[np.dtype(dtype)]
        n, rank = N.u64(), C.c_int()
        N.check(N.lib().dxg_h5_read_attribute(self._h, path.encode(), name.encode(), t, None, 0, C.byref(n), C.byref(rank)))
        if n.value == 0:
            return None
        out = np.empty(n.value, dtype=dtype)
        N.check(N.lib().dxg_h5_read_attribute(self._h, path.encode(), name.encode(), t, out.ctypes.data_as(C.c_void_p),
                                              out.size, C.byref(n), C.byref(rank)))
        return out, rank.value


def last_deflate() -> tuple[float, int]:
    """(ms, threads) of the last compressed dataset written on this thread."""
    t = C.c_int()
    ms = N.lib().dxg_h5_last_deflate_ms(C.byref(t))
    return float(ms), t.value


# ------------------------------------------------------------------ data


@dataclass
class ImageData:
    """The parts of DataContainer a project file holds.  Volumes are flat,
    x fastest, of dimensions[0] * dimensions[1] * dimensions[2] voxels."""
    dimensions: tuple[int, int, int]
    spacing: tuple[float, float, float]
    density: np.ndarray | None = None
    ct: np.ndarray | None = None
    material: np.ndarray | None = None
    organ: np.ndarray | None = None
    dose: np.ndarray | None = None
    dose_variance: np.ndarray | None = None
    dose_count: np.ndarray | None = None
    organ_names: list[str] = field(default_factory=list)
    materials: list[tuple[str, dict[int, float]]] = field(default_factory=list)
    aec: CTAECFilter | None = None

    def size(self) -> int:
        return int(np.prod(self.dimensions))


_VOLUMES = (("densityarray", "density", np.float64), ("ctarray", "ct", np.float64), ("materialarray", "material", np.uint8),
            ("organarray", "organ", np.uint8))
_DOSE_VOLUMES = (("dosearray", "dose"), ("dosevariancearray", "dose_variance"), ("doseeventcountarray", "dose_count"))



def _deg(rad) -> np.ndarray:
    return np.degrees(np.asarray(rad, dtype=np.float64))


class HDF5Wrapper:
    """HDF5Wrapper(path, FileOpenMode): WriteOver truncates, ReadOnly reads."""

    WriteOver, ReadOnly = "w", "r"

    def __init__(self, path: str, mode: str = "r"):
        try:
            self._f: H5File | None = H5File(path, mode)
        except N.NativeError:
            self._f = None

    def close(self) -> None:
        if self._f is not None:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ volumes
    def _save_array(self, path: str, a: np.ndarray, dims=None, compress: bool = True) -> None:
        a = np.ascontiguousarray(a)
        if dims is not None:
            a = a.reshape(dims[2], dims[1], dims[0])
        self._f.write_array(path, a, 6 if (compress and a.size > 128) else -1)

    def save(self, data: ImageData) -> bool:
        """HDF5Wrapper::save(DataContainer) (hdf5wrapper.cpp:384-459)."""
        if self._f is None or data is None:
            return False
        dims = tuple(int(d) for d in data.dimensions)
        self._save_array("dimensions", np.array(dims, dtype=np.uint64))
        self._save_array("spacing", np.array(data.spacing, dtype=np.float64))
        for name, attr, dtype in _VOLUMES:
            v = getattr(data, attr)
            if v is not None and np.size(v) > 0:
                self._save_array(name, np.asarray(v, dtype=dtype).ravel(), dims)
        if data.organ_names:
            self._f.write_strings("organnames", list(data.organ_names))
        if data.materials:
            self._f.write_strings("materialnames", [m[0] for m in data.materials])
            self._f.write_strings("materialcomposition", [composition_string(m[1]) for m in data.materials])
        for name, attr in _DOSE_VOLUMES:
            v = getattr(data, attr)
            if v is not None and np.size(v) > 0:
                self._save_array(name, np.asarray(v, dtype=np.float64).ravel(), dims)
        if data.aec is not None and len(data.aec.weights) > 2:
            self._save_array("aecweights", np.asarray(data.aec.weights, dtype=np.float64))
            self._save_array("aecstart", np.asarray(data.aec.start, dtype=np.float64))
            self._save_array("aecstop", np.asarray(data.aec.stop, dtype=np.float64))
        return True

    def _load(self, path: str, dtype) -> np.ndarray:
        if self._f.kind(path) != 2:
            return np.zeros(0, dtype=dtype)
        return self._f.read_array(path, dtype).ravel()

    def load(self) -> ImageData | None:
        """HDF5Wrapper::load() (hdf5wrapper.cpp:1070-1150)."""
        if self._f is None:
            return None
        dims = self._load("dimensions", np.uint64)
        if dims.size != 3:
            return None
        spacing = self._load("spacing", np.float64)
        if spacing.size != 3:
            return None
        res = ImageData(tuple(int(d) for d in dims), tuple(float(s) for s in spacing))
        n = res.size()
        material = self._load("materialarray", np.uint8)
        if material.size != n:
            return None
        res.material = material
        names = self._f.read_strings("materialnames") if self._f.kind("materialnames") == 2 else []
        comps = self._f.read_strings("materialcomposition") if self._f.kind("materialcomposition") == 2 else []
        if len(names) != len(comps):
            return None
        res.materials = [(nm, parse_composition(c)) for nm, c in zip(names, comps)]
        organ = self._load("organarray", np.uint8)
        if organ.size == n:
            res.organ = organ
            res.organ_names = self._f.read_strings("organnames") if self._f.kind("organnames") == 2 else []
        density = self._load("densityarray", np.float64)
        if density.size != n:
            return None
        res.density = density
        for name, attr in (("ctarray", "ct"),) + _DOSE_VOLUMES:
            v = self._load(name, np.float64)
            if v.size == n:
                setattr(res, attr, v)
        start, stop = self._load("aecstart", np.float64), self._load("aecstop", np.float64)
        weights = self._load("weights", np.float64)   # saved as "aecweights"
        if start.size == 3 and stop.size == 3 and weights.size > 2:
            res.aec = CTAECFilter(tuple(start), tuple(stop), weights)
        return res

    # ------------------------------------------------------------ beams
    def _next_group(self, kind: str) -> str:
        k = 1
        while self._f.kind(f"/beams/{kind}/{k}") == 1:
            k += 1
        path = f"/beams/{kind}/{k}"
        self._f.create_group(path)
        return path

    def _attrs(self, path: str, values: list[tuple[str, object, type]]) -> None:
        for name, v, dtype in values:
            self._f.write_attribute(path, name, v, dtype=dtype)

    @staticmethod
    def _tube(tube, suffix: str = "") -> list[tuple[str, object, type]]:
        f = tube.filtration
        return [(f"tube_voltage{suffix}", float(tube.voltage), np.float64),
                (f"tube_Al_filtration{suffix}", float(f.get(13, 0.0)), np.float64),
                (f"tube_Cu_filtration{suffix}", float(f.get(29, 0.0)), np.float64),
                (f"tube_Sn_filtration{suffix}", float(f.get(50, 0.0)), np.float64)]

    @staticmethod
    def _organ_aec(beam) -> list[tuple[str, object, type]]:
        o = beam.organAECFilter()
        angles = (o.startAngle(), o.stopAngle(), o.rampAngle())
        return [("use_organ_aec", np.uint64(o.useFilter()), np.uint64),
                ("compensate_weight_organ_aec", np.uint64(o.compensateOutside()), np.uint64),
                ("low_weight_organ_aec", float(o.lowWeight()), np.float64),
                ("start_stop_ramp_angle_organ_aec", np.asarray(angles, dtype=np.float64), np.float64)]

    def saveBeam(self, beam) -> bool:
        """HDF5Wrapper::save(BeamActorContainer) (hdf5wrapper.cpp:461-628, 1052-1068)."""
        if self._f is None or beam is None:
            return False
        ppe = np.uint64(beam.numberOfParticlesPerExposure)
        if isinstance(beam, DXBeam):
            path = self._next_group("DXBeams")
            self._attrs(path, [("rotation_center", np.asarray(beam.rotation_center, float), np.float64),
                               ("source_center_distance", float(beam.spd), np.float64),
                               ("primary_angle", float(_deg(beam.angles[0])), np.float64),
                               ("secondary_angle", float(_deg(beam.angles[1])), np.float64),
                               ("source_detector_distance", float(beam.sdd), np.float64),
                               ("collimation_half_angles", _deg(beam.half_angles), np.float64),
                               ("DAPvalue", float(beam.DAPvalue()), np.float64),
                               ("number_of_exposures", np.uint64(beam.n_exposures), np.uint64),
                               ("particles_per_exposure", ppe, np.uint64),
                               ("tube_voltage", float(beam.tube.voltage), np.float64),
                               ("tube_anode_angle", float(beam.tube.anode_angle_deg), np.float64)]
                        + self._tube(beam.tube)[1:])
        elif isinstance(beam, CBCTBeam):
            path = self._next_group("CBCTBeams")
            self._attrs(path, [("isocenter", np.asarray(beam.isocenter, float), np.float64),
                               ("source_detector_distance", float(beam.sdd), np.float64),
                               ("start_angle", float(beam.start_angle), np.float64),
                               ("stop_angle", float(beam.stop_angle), np.float64),
                               ("step_angle", float(beam.step_angle), np.float64),
                               ("collimation_half_angles", _deg(beam.half_angles), np.float64),
                               ("DAPvalue", float(beam.DAPvalue()), np.float64),
                               ("particles_per_exposure", ppe, np.uint64),
                               ("tube_voltage", float(beam.tube.voltage), np.float64),
                               ("tube_anode_angle", float(beam.tube.anode_angle_deg), np.float64)]
                        + self._tube(beam.tube)[1:])
        elif isinstance(beam, CTSequentialBeam):
            path = self._next_group("CTSequentialBeams")
            self._attrs(path, [("start_position", np.asarray(beam.position, float), np.float64),
                               ("scan_normal", np.asarray(beam.scan_normal, float), np.float64),
                               ("scan_field_view", float(beam.fov), np.float64),
                               ("source_detector_distance", float(beam.sdd), np.float64),
                               ("collimation", float(beam.collimation), np.float64),
                               ("start_angle", float(_deg(beam.start_angle)), np.float64),
                               ("step_angle", float(_deg(beam.step_angle)), np.float64),
                               ("slice_spacing", float(beam.slice_spacing), np.float64),
                               ("CTDIw", float(beam.CTDIw()), np.float64),
                               ("CTDIdiameter", float(beam.CTDIdiameter()), np.float64),
                               ("tube_voltage", float(beam.tube.voltage), np.float64),
                               ("tube_anode_angle", float(beam.tube.anode_angle_deg), np.float64)]
                        + self._tube(beam.tube)[1:]
                        + [("particles_per_exposure", ppe, np.uint64),
                           ("number_of_slices", np.uint64(beam.slices), np.uint64)])
        elif isinstance(beam, CTSpiralDualEnergyBeam):
            path = self._next_group("CTSpiralDualEnergyBeams")
            self._attrs(path, [("start_position", np.asarray(beam.start, float), np.float64),
                               ("stop_position", np.asarray(beam.stop, float), np.float64),
                               ("scan_field_viewA", float(beam.fov), np.float64),
                               ("scan_field_viewB", float(beam.fov_b), np.float64),
                               ("source_detector_distance", float(beam.sdd), np.float64),
                               ("collimation", float(beam.collimation), np.float64),
                               ("start_angle", float(_deg(beam.start_angle)), np.float64),
                               ("step_angle", float(_deg(beam.step_angle)), np.float64),
                               ("pitch", float(beam.pitch), np.float64),
                               ("CTDIvol", float(beam.CTDIvol()), np.float64),
                               ("CTDIdiameter", float(beam.CTDIdiameter()), np.float64)]
                        + self._tube(beam.tubeB, "B")
                        + [("tube_anode_angle", float(beam.tube.anode_angle_deg), np.float64)]
                        + self._tube(beam.tube, "A")
                        + [("tube_masA", float(beam.mas_a), np.float64), ("tube_masB", float(beam.mas_b), np.float64)]
                        + self._organ_aec(beam) + [("particles_per_exposure", ppe, np.uint64)])
        elif isinstance(beam, CTSpiralBeam):
            path = self._next_group("CTSpiralBeams")
            self._attrs(path, [("start_position", np.asarray(beam.start, float), np.float64),
                               ("stop_position", np.asarray(beam.stop, float), np.float64),
                               ("scan_field_view", float(beam.fov), np.float64),
                               ("source_detector_distance", float(beam.sdd), np.float64),
                               ("collimation", float(beam.collimation), np.float64),
                               ("start_angle", float(_deg(beam.start_angle)), np.float64),
                               ("step_angle", float(_deg(beam.step_angle)), np.float64),
                               ("pitch", float(beam.pitch), np.float64),
                               ("CTDIvol", float(beam.CTDIvol()), np.float64),
                               ("CTDIdiameter", float(beam.CTDIdiameter()), np.float64),
                               ("tube_voltage", float(beam.tube.voltage), np.float64),
                               ("tube_anode_angle", float(beam.tube.anode_angle_deg), np.float64)]
                        + self._tube(beam.tube)[1:] + self._organ_aec(beam)
                        + [("particles_per_exposure", ppe, np.uint64)])
        else:
            return False
        return True

    def _get(self, path: str, name: str, n: int = 1, dtype=np.float64):
        r = self._f.read_attribute(path, name, dtype)
        if r is None or r[0].size != n:
            return None
        return r[0] if n > 1 else r[0][0]

    def _load_tube(self, path: str, tube, suffix: str = "") -> None:
        v = self._get(path, f"tube_voltage{suffix}")
        if v is not None:
            tube.voltage = float(v)
        for z, sym in ((13, "Al"), (29, "Cu"), (50, "Sn")):
            f = self._get(path, f"tube_{sym}_filtration{suffix}")
            if f is not None:
                tube.filtration[z] = float(f)   # addTubeFiltrationMaterial

    def _load_organ_aec(self, path: str, beam) -> None:
        o = beam.organAECFilter()
        u = self._get(path, "use_organ_aec", dtype=np.uint64)
        if u is not None:
            o.setUseFilter(int(u) == 1)
        a = self._get(path, "start_stop_ramp_angle_organ_aec", 3)
        if a is not None:
            o.setStartAngle(float(a[0]))
            o.setStopAngle(float(a[1]))
            o.setRampAngle(float(a[2]))
        c = self._get(path, "compensate_weight_organ_aec", dtype=np.uint64)
        if c is not None:
            o.setCompensateOutside(int(c) == 1)
        w = self._get(path, "low_weight_organ_aec")
        if w is not None:
            o.setLowWeightFactor(float(w))

    def _load_beam(self, kind: str, path: str):
        g = lambda name, n=1, dtype=np.float64: self._get(path, name, n, dtype)   # noqa: E731
        if kind == "DXBeams":
            b = DXBeam()
            if (v := g("rotation_center", 3)) is not None:
                b.rotation_center = tuple(v)
            if (v := g("source_center_distance")) is not None:
                b.spd = float(v)
            if (v := g("primary_angle")) is not None:
                b.angles = (math.radians(v), b.angles[1])
            if (v := g("secondary_angle")) is not None:
                b.angles = (b.angles[0], math.radians(v))
            if (v := g("source_detector_distance")) is not None:
                b.sdd = float(v)
            if (v := g("collimation_half_angles", 2)) is not None:
                b.setCollimationHalfAngles(np.radians(v))
            if (v := g("DAPvalue")) is not None:
                b.setDAPvalue(v)
            if (v := g("number_of_exposures", dtype=np.uint64)) is not None:
                b.setNumberOfExposures(int(v))
        elif kind == "CBCTBeams":
            b = CBCTBeam()
            if (v := g("isocenter", 3)) is not None:
                b.isocenter = tuple(v)
            if (v := g("source_detector_distance")) is not None:
                b.sdd = float(v)
            for name in ("start_angle", "stop_angle", "step_angle"):
                if (v := g(name)) is not None:
                    setattr(b, name, float(v))
            if (v := g("collimation_half_angles", 2)) is not None:
                b.half_angles = tuple(np.radians(v))
            if (v := g("DAPvalue")) is not None:
                b.setDAPvalue(v)
        elif kind == "CTSequentialBeams":
            b = CTSequentialBeam()
            if (v := g("position", 3)) is not None:        # saved as "start_position"
                b.position = tuple(v)
            if (v := g("scan_normal", 3)) is not None:
                b.scan_normal = tuple(v)
            if (v := g("slice_spacing")) is not None:
                b.setSliceSpacing(v)
            if (v := g("number_of_slices", dtype=np.uint64)) is not None:
                b.setNumberOfSlices(int(v))
            if (v := g("scan_field_view")) is not None:
                b.setScanFieldOfView(v)
            if (v := g("CTDIw")) is not None:
                b.setCTDIw(v)
            if (v := g("CTDIDiameter")) is not None:       # saved as "CTDIdiameter"
                b.setCTDIdiameter(v)
        else:   # CTSpiralBeams / CTDualEnergySpiralBeams
            dual = kind == "CTDualEnergySpiralBeams"
            b = CTSpiralDualEnergyBeam() if dual else CTSpiralBeam()
            if (v := g("start_position", 3)) is not None:
                b.start = tuple(v)
            if (v := g("stop_position", 3)) is not None:
                b.stop = tuple(v)
            if dual:
                if (v := g("scan_field_viewA")) is not None:
                    b.setScanFieldOfViewA(v)
                if (v := g("scan_field_viewB")) is not None:
                    b.setScanFieldOfViewB(v)
            elif (v := g("scan_field_view")) is not None:
                b.setScanFieldOfView(v)
            if (v := g("pitch")) is not None:
                b.setPitch(v)
            if (v := g("CTDIvol")) is not None:
                b.setCTDIvol(v)
            if (v := g("CTDIDiameter")) is not None:       # saved as "CTDIdiameter"
                b.setCTDIdiameter(v)
            self._load_organ_aec(path, b)
        if kind in ("CTSequentialBeams", "CTSpiralBeams", "CTDualEnergySpiralBeams"):
            if (v := g("source_detector_distance")) is not None:
                b.setSourceDetectorDistance(v)
            if (v := g("collimation")) is not None:
                b.setCollimation(v)
            if (v := g("start_angle")) is not None:
                b.setStartAngleDeg(v)
            if (v := g("step_angle")) is not None:
                b.setStepAngleDeg(v)
        if (v := g("particles_per_exposure", dtype=np.uint64)) is not None:
            b.setNumberOfParticlesPerExposure(int(v))
        if (v := g("tube_anode_angle")) is not None:
            b.tube.anode_angle_deg = float(v)
            if kind == "CTDualEnergySpiralBeams":
                b.tubeB.anode_angle_deg = float(v)
        if kind == "CTDualEnergySpiralBeams":
            self._load_tube(path, b.tube, "A")
            self._load_tube(path, b.tubeB, "B")
            if (v := g("tube_masA")) is not None:
                b.setRelativeMasTubeA(v)
            if (v := g("tube_masB")) is not None:
                b.setRelativeMasTubeB(v)
        else:
            self._load_tube(path, b.tube)
        return b

    def loadBeams(self) -> list:
        """HDF5Wrapper::loadBeams() (hdf5wrapper.cpp:1152-1207)."""
        res = []
        if self._f is None or self._f.kind("beams") != 1:
            return res
        for kind in ("DXBeams", "CTSpiralBeams", "CTDualEnergySpiralBeams", "CBCTBeams", "CTSequentialBeams"):
            k = 1
            while self._f.kind(f"beams/{kind}/{k}") == 1:
                res.append(self._load_beam(kind, f"beams/{kind}/{k}"))
                k += 1
        return res


# ------------------------------------------------------------------ timing


def synthetic_project(dims, seed: int = 1) -> ImageData:
    """A CT-like project of nx x ny x nz voxels: elliptic body with lungs and
    spine, five materials, organ labels, and a dose with 3 % Monte Carlo
    noise, its variance and event counts."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.linspace(-1, 1, nz), np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    body = (x / 0.9) ** 2 + (y / 0.6) ** 2 <= 1.0
    lung = body & (((x - 0.35) / 0.25) ** 2 + (y / 0.35) ** 2 <= 1.0) | body & (((x + 0.35) / 0.25) ** 2 + (y / 0.35) ** 2 <= 1.0)
    bone = body & ((x / 0.08) ** 2 + ((y + 0.4) / 0.08) ** 2 <= 1.0)
    material = np.where(body, 2, 0).astype(np.uint8)
    material[lung] = 1
    material[bone] = 4
    density = np.choose(material, [0.0012, 0.26, 1.0, 1.05, 1.85]) * (1 + 0.01 * rng.standard_normal(material.shape))
    organ = (material * 3 + (z > 0)).astype(np.uint8)
    dose = np.exp(-2.0 * (x ** 2 + y ** 2)) * (1.0 + 0.03 * rng.standard_normal(material.shape))
    names = ["Air, Dry (near sea level)", "Lung", "Tissue, Soft (ICRP)", "Muscle, Skeletal", "Bone, Cortical (ICRP)"]
    return ImageData(tuple(dims), (0.1, 0.1, 0.1), density=density.ravel(), ct=(density.ravel() - 1.0) * 1000.0,
                     material=material.ravel(), organ=organ.ravel(), dose=dose.ravel(),
                     dose_variance=(0.03 * dose.ravel()) ** 2, dose_count=np.rint(1e4 * dose.ravel()),
                     organ_names=[f"organ {i}" for i in range(16)],
                     materials=[(n, {1: 0.1, 8: 0.9}) for n in names])


def main() -> None:
    import argparse
    import os
    import tempfile
    import time

    ap = argparse.ArgumentParser(description="time saving and loading an OpenDXMC project file")
    ap.add_argument("--dims", type=int, nargs=3, default=(256, 256, 200))
    ap.add_argument("--dir", default=tempfile.gettempdir())
    a = ap.parse_args()
    data = synthetic_project(a.dims)
    raw = sum(np.asarray(getattr(data, f)).nbytes for f in ("density", "ct", "material", "organ", "dose", "dose_variance",
                                                              "dose_count"))
    path = os.path.join(a.dir, f"dxg_project_{os.getpid()}.h5")
    try:
        for label, env in (("threads", "0"), ("hdf5 filter", "1")):
            os.environ["DXG_H5_LIBRARY_DEFLATE"] = env
            t0 = time.perf_counter()
            with HDF5Wrapper(path, HDF5Wrapper.WriteOver) as w:
                w.save(data)
            t_save = time.perf_counter() - t0
            size = os.path.getsize(path)
            t0 = time.perf_counter()
            with HDF5Wrapper(path, HDF5Wrapper.ReadOnly) as r:
                back = r.load()
            t_load = time.perf_counter() - t0
            assert back is not None and np.array_equal(back.density, data.density)
            print(f"{label}: save {t_save:.2f} s ({raw / t_save / 1e6:.0f} MB/s of arrays), file {size / 1e6:.0f} MB "
                  f"of {raw / 1e6:.0f} MB, load {t_load:.2f} s", flush=True)
    finally:
        os.environ.pop("DXG_H5_LIBRARY_DEFLATE", None)
        if os.path.exists(path):
            os.remove(path)


if __name__ == "__main__":
    main()
