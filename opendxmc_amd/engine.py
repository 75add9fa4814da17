"""Python mirror of the DXMClib call surface OpenDXMC uses for the voxel dose
path (simulationpipeline.cpp:124-235, dxmc_specialization.hpp:37-77), over the
C-ABI of libdxmc_hip.so.  Same names and argument meaning as the reference:

    world = World(); vgrid = world.addItem(AAVoxelGrid(correction=1))
    vgrid.setData(dims, density, material, materials); vgrid.setSpacing(spacing)
    world.build()
    transport = Transport(); transport(world, beam, progress, True)
    vgrid.doseScored()  -> (dose mGy, variance mGy^2, number of events) per voxel

Failures surface as exceptions carrying the C status code; there is no CPU
fallback anywhere.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


class _Keep:
    def __init__(self):
        self.items = []

    def keep(self, x):
        self.items.append(x)
        return x


# --------------------------------------------------------------------- material
class Material:
    """Material<5>::byWeight (simulationpipeline.cpp:136): element mass fractions."""

    def __init__(self, composition: dict[int, float]):
        self.composition = {int(z): float(w) for z, w in composition.items()}

    @staticmethod
    def byWeight(composition: dict[int, float]) -> "Material | None":
        """Returns None when the material cannot be built (std::optional)."""
        m = Material(composition)
        try:
            m.attenuationValues(np.array([60.0]))
        except N.NativeError:
            return None
        return m

    def _struct(self, keep: _Keep) -> N.Material:
        z = keep.keep(np.array(list(self.composition.keys()), dtype=np.int32))
        w = keep.keep(np.array(list(self.composition.values()), dtype=np.float64))
        return N.Material(N.ptr(z, C.c_int32), N.ptr(w, C.c_double), len(z))

    def attenuationValues(self, energy) -> np.ndarray:
        """(n, 4): photoelectric, incoherent, coherent, total mass attenuation (cm2/g)."""
        k = _Keep()
        e = np.ascontiguousarray(np.atleast_1d(energy), dtype=np.float64)
        out = np.zeros((len(e), 4))
        m = self._struct(k)
        N.check(N.lib().dxg_material_attenuation(C.byref(m), N.ptr(e, C.c_double), len(e), N.ptr(out, C.c_double)))
        return out


# ------------------------------------------------------------- progress / grid
class TransportProgress:
    """TransportProgress: thread-safe progress and cancellation."""

    def __init__(self):
        h = C.c_void_p()
        N.check(N.lib().dxg_progress_create(C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().dxg_progress_destroy(self._h)
            self._h = None

    def setStopSimulation(self) -> None:
        N.lib().dxg_progress_request_stop(self._h)

    def continueSimulation(self) -> bool:
        return bool(N.lib().dxg_progress_continue(self._h))

    def progress(self) -> tuple[int, int]:
        n, t = N.u64(), N.u64()
        N.lib().dxg_progress_get(self._h, C.byref(n), C.byref(t))
        return n.value, t.value

    def message(self) -> str:
        buf = C.create_string_buffer(256)
        N.lib().dxg_progress_message(self._h, buf, 256)
        return buf.value.decode()


class AAVoxelGrid:
    """AAVoxelGrid<5, CORRECTION, 255>; the correction level is fixed at creation."""

    def __init__(self, correction: int = 1):
        h = C.c_void_p()
        N.check(N.lib().dxg_world_create(C.byref(h), int(correction)))
        self._h = h
        self.correction = correction
        self.dims = (0, 0, 0)
        self.spacing = (0.0, 0.0, 0.0)
        self.density = None
        self.material = None

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().dxg_world_destroy(self._h)
            self._h = None

    def setData(self, dims, density, material, materials: list[Material]) -> None:
        k = _Keep()
        d = np.ascontiguousarray(dims, dtype=np.uint64)
        dens = np.ascontiguousarray(density, dtype=np.float64).ravel()
        mat = np.ascontiguousarray(material, dtype=np.uint8).ravel()
        n = int(np.prod(d))
        if dens.size != n or mat.size != n:
            raise ValueError("density/material size does not match dims")
        ms = (N.Material * len(materials))(*[m._struct(k) for m in materials])
        N.check(N.lib().dxg_grid_set_data(self._h, N.ptr(d, N.u64), N.ptr(dens, C.c_double),
                                          N.ptr(mat, C.c_uint8), ms, len(materials)))
        self.dims = tuple(int(x) for x in d)
        self.density = dens
        self.material = mat

    def setSpacing(self, spacing) -> None:
        sp = np.ascontiguousarray(spacing, dtype=np.float64)
        N.check(N.lib().dxg_grid_set_spacing(self._h, N.ptr(sp, C.c_double)))
        self.spacing = tuple(float(x) for x in sp)

    def size(self) -> int:
        return int(N.lib().dxg_grid_size(self._h))

    def build(self) -> None:
        N.check(N.lib().dxg_world_build(self._h))

    def clearScores(self) -> None:
        N.check(N.lib().dxg_world_clear_scores(self._h))

    def histories(self) -> int:
        return int(N.lib().dxg_world_histories(self._h))

    def energyScored(self):
        """(sum w dE keV, sum (w dE)^2 keV^2, events) per voxel."""
        n = self.size()
        e, e2, ev = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.uint64)
        N.check(N.lib().dxg_read_energy_scores(self._h, N.ptr(e, C.c_double), N.ptr(e2, C.c_double), N.ptr(ev, N.u64)))
        return e, e2, ev

    def doseScored(self):
        """(dose mGy, variance mGy^2, numberOfEvents) per voxel (DoseScore)."""
        n = self.size()
        d, v, ev = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.uint64)
        N.check(N.lib().dxg_read_dose_scores(self._h, N.ptr(d, C.c_double), N.ptr(v, C.c_double), N.ptr(ev, N.u64)))
        return d, v, ev

    def counters(self) -> dict:
        c = N.Counters()
        N.check(N.lib().dxg_get_counters(self._h, C.byref(c)))
        return c.as_dict()

    def traverse(self, pos, dirs, steps):
        p = np.ascontiguousarray(pos, dtype=np.float32)
        d = np.ascontiguousarray(dirs, dtype=np.float32)
        s = np.ascontiguousarray(steps, dtype=np.float32)
        nr, ns = s.shape
        idx = np.zeros((nr, ns), dtype=np.int64)
        mm = np.zeros((nr, ns), dtype=np.int32)
        N.check(N.lib().dxg_debug_traverse(self._h, nr, N.ptr(p, C.c_float), N.ptr(d, C.c_float), ns,
                                           N.ptr(s, C.c_float), N.ptr(idx, C.c_int64), N.ptr(mm, C.c_int32)))
        return idx, mm


class World:
    """World<AAVoxelGrid>: owns its one voxel-grid item."""

    def __init__(self):
        self.items: list[AAVoxelGrid] = []

    def addItem(self, item: AAVoxelGrid) -> AAVoxelGrid:
        if self.items:
            raise ValueError("this World holds one AAVoxelGrid")
        self.items.append(item)
        return item

    def build(self) -> None:
        for it in self.items:
            it.build()


# ---------------------------------------------------------------------- beams
@dataclass
class Tube:
    """Tube: voltage (kV), anode angle (deg) and filtration {Z: mm}."""
    voltage: float = 120.0
    anode_angle_deg: float = 12.0
    filtration: dict = field(default_factory=dict)
    energy_step: float = 0.5

    def _struct(self, keep: _Keep):
        z = keep.keep(np.array(list(self.filtration.keys()) or [13], dtype=np.int32))
        mm = keep.keep(np.array(list(self.filtration.values()) or [0.0], dtype=np.float64))
        return N.Tube(self.voltage, self.anode_angle_deg, N.ptr(z, C.c_int32), N.ptr(mm, C.c_double),
                      len(self.filtration), self.energy_step)

    def getSpecter(self) -> tuple[np.ndarray, np.ndarray]:
        # the spectrum depends on the tube's settings only: computed once per
        # setting (every Transport call flattens the beam's spectra again)
        key = (float(self.voltage), float(self.anode_angle_deg),
               tuple(sorted((int(z), float(mm)) for z, mm in self.filtration.items())), float(self.energy_step))
        cached = self.__dict__.get("_spectrum_cache")
        if cached is None or cached[0] != key:
            k = _Keep()
            t = self._struct(k)
            n = N.u64()
            N.check(N.lib().dxg_tube_spectrum(C.byref(t), None, None, 0, C.byref(n)))
            e, w = np.zeros(n.value), np.zeros(n.value)
            N.check(N.lib().dxg_tube_spectrum(C.byref(t), N.ptr(e, C.c_double), N.ptr(w, C.c_double), n.value, C.byref(n)))
            cached = (key, e, w)
            self.__dict__["_spectrum_cache"] = cached
        return cached[1].copy(), cached[2].copy()

    def halfValueLayerAl(self) -> float:
        k = _Keep()
        t = self._struct(k)
        v = C.c_double()
        N.check(N.lib().dxg_tube_hvl_mm_al(C.byref(t), C.byref(v)))
        return v.value

    def meanSpecterEnergy(self) -> float:
        k = _Keep()
        t = self._struct(k)
        v = C.c_double()
        N.check(N.lib().dxg_tube_mean_energy(C.byref(t), C.byref(v)))
        return v.value


class BowtieFilter:
    """BowtieFilter(vector<pair<angle, weight>>) (bowtiefilterreader.cpp:92)."""

    def __init__(self, pairs):
        pairs = list(pairs)
        self.angle = np.array([a for a, _ in pairs], dtype=np.float64)
        self.weight = np.array([w for _, w in pairs], dtype=np.float64)


@dataclass
class CTAECFilter:
    """CTAECFilter(start, stop, weights) (datacontainer.cpp:59)."""
    start: tuple = (0.0, 0.0, 0.0)
    stop: tuple = (0.0, 0.0, 1.0)
    weights: tuple = ()

    def isEmpty(self) -> bool:
        return len(self.weights) == 0

    @staticmethod
    def fromWaterEquivalentDiameter(dims, spacing, image, is_ct: bool = True) -> "CTAECFilter":
        """DataContainer::calculateAECfilterFromWaterEquivalentDiameter
        (datacontainer.cpp:42-100): weight exp(0.2 D_w) per z slice of a CT
        (HU) or density image, from (0, 0, -L) to (0, 0, L)."""
        d = np.ascontiguousarray(dims, dtype=np.uint64)
        sp = np.ascontiguousarray(spacing, dtype=np.float64)
        img = np.ascontiguousarray(image, dtype=np.float64).ravel()
        if img.size != int(np.prod(d)):
            raise ValueError("image size does not match dims")
        w = np.zeros(int(d[2]))
        start, stop = (C.c_double * 3)(), (C.c_double * 3)()
        N.check(N.lib().dxg_aec_from_wed(N.ptr(d, N.u64), N.ptr(sp, C.c_double), N.ptr(img, C.c_double), int(bool(is_ct)),
                                         N.ptr(w, C.c_double), start, stop))
        return CTAECFilter(tuple(start), tuple(stop), tuple(w))


class CTOrganAECFilter:
    """The organ AEC filter of CT spiral beams (organAECFilter(); settings at
    beamsettingsmodel.cpp:349-437, stored at hdf5wrapper.cpp:576-580): a low
    tube-current weight over a gantry-angle arc, linear ramps beside it, and
    optionally a raised weight elsewhere that keeps the mean over a turn at 1."""

    def __init__(self):
        self._use = False
        self._compensate = False
        self._start = 0.0
        self._stop = 0.0
        self._ramp = 0.0
        self._low = 1.0

    def setUseFilter(self, on: bool) -> None: self._use = bool(on)
    def useFilter(self) -> bool: return self._use
    def setCompensateOutside(self, on: bool) -> None: self._compensate = bool(on)
    def compensateOutside(self) -> bool: return self._compensate
    def setStartAngle(self, rad: float) -> None: self._start = float(rad)
    def setStopAngle(self, rad: float) -> None: self._stop = float(rad)
    def setRampAngle(self, rad: float) -> None: self._ramp = float(rad)
    def startAngle(self) -> float: return self._start
    def stopAngle(self) -> float: return self._stop
    def rampAngle(self) -> float: return self._ramp
    def setStartAngleDeg(self, deg: float) -> None: self._start = math.radians(deg)
    def setStopAngleDeg(self, deg: float) -> None: self._stop = math.radians(deg)
    def setRampAngleDeg(self, deg: float) -> None: self._ramp = math.radians(deg)
    def startAngleDeg(self) -> float: return math.degrees(self._start)
    def stopAngleDeg(self) -> float: return math.degrees(self._stop)
    def rampAngleDeg(self) -> float: return math.degrees(self._ramp)
    def setLowWeightFactor(self, w: float) -> None: self._low = float(w)
    def lowWeight(self) -> float: return self._low

    def _struct(self) -> "N.OrganAEC":
        return N.OrganAEC(int(self._use), int(self._compensate), self._start, self._stop, self._ramp, self._low)

    def maxWeight(self) -> float:
        v = C.c_double()
        f = self._struct()
        N.check(N.lib().dxg_organ_aec_max_weight(C.byref(f), C.byref(v)))
        return v.value

    def weight(self, angle: float) -> float:
        """The weight at gantry angle `angle` (rad)."""
        v = C.c_double()
        f = self._struct()
        N.check(N.lib().dxg_organ_aec_weight(C.byref(f), float(angle), C.byref(v)))
        return v.value


class _BeamBase:
    numberOfParticlesPerExposure = 1_000_000

    def setNumberOfParticlesPerExposure(self, n: int) -> None:
        self.numberOfParticlesPerExposure = int(n)

    def numberOfExposures(self) -> int:
        return len(self.exposures())

    def numberOfParticles(self) -> int:
        return int(self.exposures()["n_particles"].sum())

    def _collect(self, builder, params) -> np.ndarray:
        n = N.u64()
        N.check(builder(C.byref(params), None, 0, C.byref(n)))
        arr = np.zeros(n.value, dtype=N.EXPOSURE_DTYPE)
        N.check(builder(C.byref(params), arr.ctypes.data_as(C.POINTER(N.Exposure)), n.value, C.byref(n)))
        return arr

    def spectra(self) -> list[tuple[np.ndarray, np.ndarray]]:
        return [self.tube.getSpecter()]

    def bowties(self) -> list[BowtieFilter]:
        return []

    def calibrationFactor(self) -> float:
        """mGy per (dose summed over a run's histories) for the beam's stated
        output (DAP, air KERMA or CTDI); 0 when none is set."""
        return 0.0


def _calibrate(fn, beam, value: float) -> float:
    k = _Keep()
    b = beam_struct(beam, k, calibrated=False)
    f = C.c_double()
    N.check(fn(C.byref(b), float(value), C.byref(f)))
    return f.value


class PencilBeam(_BeamBase):
    def __init__(self, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0), energy: float = 60.0):
        self.position = tuple(position)
        self.direction = tuple(direction)
        self.energy = energy
        self.n_exposures = 1

    def setEnergy(self, e: float) -> None:
        self.energy = min(max(float(e), 20.0), 150.0)   # beamsettingsmodel.cpp:664

    def setNumberOfExposures(self, n: int) -> None:
        self.n_exposures = int(n)

    def exposures(self) -> np.ndarray:
        p = N.PencilParams(tuple(self.position), tuple(self.direction), self.n_exposures, self.numberOfParticlesPerExposure)
        return self._collect(N.lib().dxg_pencil_exposures, p)

    def spectra(self):
        return [(np.array([float(self.energy)]), np.array([1.0]))]

    air_kerma = 0.0

    def setAirKerma(self, mgy: float) -> None:
        """Air KERMA (mGy) over a 1 cm2 cross-section of the beam (beamsettingsmodel.cpp:673-681)."""
        self.air_kerma = float(mgy)

    def airKerma(self) -> float:
        return self.air_kerma

    def calibrationFactor(self) -> float:
        return _calibrate(N.lib().dxg_calibration_air_kerma, self, self.air_kerma) if self.air_kerma > 0 else 0.0


class DXBeam(_BeamBase):
    """dxmc::DXBeam with OpenDXMC's geometry wrapper (dxmc_specialization.cpp:21-90)."""

    def __init__(self, filtration=None):
        self.tube = Tube(voltage=70.0, filtration=dict(filtration or {}))
        self.rotation_center = (0.0, 0.0, 0.0)
        self.spd = 100.0
        self.sdd = 100.0
        self.angles = (0.0, 0.0)
        self.n_exposures = 1
        self.half_angles = (0.0, 0.0)
        self.setCollimation((20.0, 20.0))

    def setCollimation(self, coll) -> None:
        # tan(0.5 |coll| / SDD) as in dxmc_specialization.cpp:46-52
        self.half_angles = (math.tan(0.5 * abs(coll[0]) / self.sdd), math.tan(0.5 * abs(coll[1]) / self.sdd))

    def setCollimationHalfAngles(self, ang) -> None:
        self.half_angles = (abs(ang[0]), abs(ang[1]))

    def setNumberOfExposures(self, n: int) -> None:
        self.n_exposures = int(n)

    def _geometry(self):
        def rot(v, axis, a):
            v, k = np.asarray(v, float), np.asarray(axis, float)
            return v * math.cos(a) + np.cross(k, v) * math.sin(a) + k * np.dot(k, v) * (1 - math.cos(a))
        c0, c1 = np.array([0.0, 0.0, 1.0]), np.array([-1.0, 0.0, 0.0])
        c0, c1 = rot(c0, (0, 0, 1), self.angles[0]), rot(c1, (0, 0, 1), self.angles[0])
        c0, c1 = rot(c0, (-1, 0, 0), self.angles[1]), rot(c1, (-1, 0, 0), self.angles[1])
        d = np.cross(c0, c1)
        pos = np.asarray(self.rotation_center) - self.spd * d
        return pos, c0, c1

    dap = 0.0

    def setDAPvalue(self, mgycm2: float) -> None:
        """Dose-area product (mGy cm2) of the beam (beamsettingsmodel.cpp:574-583)."""
        self.dap = float(mgycm2)

    def DAPvalue(self) -> float:
        return self.dap

    def calibrationFactor(self) -> float:
        return _calibrate(N.lib().dxg_calibration_dap, self, self.dap) if self.dap > 0 else 0.0

    def exposures(self) -> np.ndarray:
        pos, c0, c1 = self._geometry()
        p = N.DXParams(tuple(pos), tuple(np.concatenate([c0, c1])), tuple(self.half_angles), self.n_exposures,
                       self.numberOfParticlesPerExposure)
        return self._collect(N.lib().dxg_dx_exposures, p)


class CTSpiralBeam(_BeamBase):
    """CTSpiralBeam<false> with the reference's creation defaults (beamsettingsmodel.cpp:1150-1170)."""

    dual = False

    def __init__(self, start=(0.0, 0.0, -20.0), stop=(0.0, 0.0, 20.0), filtration=None):
        self.start = tuple(start)
        self.stop = tuple(stop)
        self.tube = Tube(voltage=120.0, filtration=dict(filtration if filtration is not None else {13: 9.0}))
        self.sdd = 119.0
        self.collimation = 3.84
        self.step_angle = math.radians(5.0)
        self.start_angle = 0.0
        self.pitch = 1.0
        self.fov = 50.0
        self.bowtie: BowtieFilter | None = None
        self.aec: CTAECFilter | None = None
        self._organ_aec = CTOrganAECFilter()

    def organAECFilter(self) -> CTOrganAECFilter:
        return self._organ_aec

    def setSourceDetectorDistance(self, d): self.sdd = float(d)
    def setCollimation(self, c): self.collimation = float(c)
    def setStepAngleDeg(self, a): self.step_angle = math.radians(a)
    def setStartAngleDeg(self, a): self.start_angle = math.radians(a)
    def setPitch(self, p): self.pitch = float(p)
    def setScanFieldOfView(self, f): self.fov = float(f)
    def setBowtieFilter(self, b): self.bowtie = b
    def setAECFilter(self, a): self.aec = a
    def setTubeVoltage(self, kv): self.tube.voltage = float(kv)
    def setTubeAnodeAngleDeg(self, a): self.tube.anode_angle_deg = float(a)
    def addTubeFiltrationMaterial(self, z, mm): self.tube.filtration[int(z)] = float(mm)

    def _params(self, keep: _Keep):
        aec = N.AEC((0.0,) * 3, (0.0, 0.0, 1.0), None, 0)
        if self.aec is not None and len(self.aec.weights):
            w = keep.keep(np.ascontiguousarray(self.aec.weights, dtype=np.float64))
            aec = N.AEC(tuple(self.aec.start), tuple(self.aec.stop), N.ptr(w, C.c_double), len(w))
        return N.CTSpiralParams(
            tuple(self.start), tuple(self.stop), self.start_angle, self.step_angle, self.pitch, self.collimation,
            self.sdd, self.fov, self.numberOfParticlesPerExposure, int(self.dual), int(bool(self.bowtie_list())),
            getattr(self, "fov_b", self.fov), getattr(self, "tube_b_offset", 0.0),
            getattr(self, "mas_a", 1.0), getattr(self, "mas_b", 1.0), aec, self._organ_aec._struct())

    def bowtie_list(self):
        return [self.bowtie] if self.bowtie is not None else []

    def bowties(self):
        return self.bowtie_list()

    def exposures(self) -> np.ndarray:
        k = _Keep()
        return self._collect(N.lib().dxg_ct_spiral_exposures, self._params(k))

    ctdi_vol = 0.0
    ctdi_diameter = 32.0

    def setCTDIvol(self, mgy: float) -> None:
        """CTDIvol (mGy) the scan is calibrated to (beamsettingsmodel.cpp:1276-1284)."""
        self.ctdi_vol = float(mgy)

    def CTDIvol(self) -> float:
        return self.ctdi_vol

    def setCTDIdiameter(self, cm: float) -> None:
        """Diameter (cm) of the PMMA CTDI phantom (beamsettingsmodel.cpp:1286-1295)."""
        self.ctdi_diameter = float(cm)

    def CTDIdiameter(self) -> float:
        return self.ctdi_diameter

    def _rotation_tubes(self):
        """(tube, spectrum, bowtie, start angle, fov, relative mAs) per tube."""
        return [(self.tube, 0, self.bowtie, self.start_angle, self.fov, getattr(self, "mas_a", 1.0))]

    def calibrationFactor(self) -> float:
        if self.ctdi_vol <= 0:
            return 0.0
        return _ctdi_factor(self, self.ctdi_vol * self.pitch)


class CTSpiralDualEnergyBeam(CTSpiralBeam):
    """CTSpiralDualEnergyBeam<false>: tubes A and B, exposures interleaved A, B."""

    dual = True

    def __init__(self, start=(0.0, 0.0, -20.0), stop=(0.0, 0.0, 20.0), filtration=None):
        super().__init__(start, stop, filtration)
        self.tubeB = Tube(voltage=140.0, filtration=dict(self.tube.filtration))
        self.fov_b = 33.0
        self.tube_b_offset = math.radians(90.0)
        self.mas_a = 1.0
        self.mas_b = 1.0
        self.bowtieB: BowtieFilter | None = None

    def setTubeAVoltage(self, kv): self.tube.voltage = float(kv)
    def setTubeBVoltage(self, kv): self.tubeB.voltage = float(kv)
    def setScanFieldOfViewA(self, f): self.fov = float(f)
    def setScanFieldOfViewB(self, f): self.fov_b = float(f)
    def setTubeBoffsetAngleDeg(self, a): self.tube_b_offset = math.radians(a)
    def setRelativeMasTubeA(self, m): self.mas_a = float(m)
    def setRelativeMasTubeB(self, m): self.mas_b = float(m)
    def setBowtieFilterA(self, b): self.bowtie = b
    def setBowtieFilterB(self, b): self.bowtieB = b

    def bowtie_list(self):
        if self.bowtie is None and self.bowtieB is None:
            return []
        a = self.bowtie or self.bowtieB
        return [a, self.bowtieB or a]

    def spectra(self):
        return [self.tube.getSpecter(), self.tubeB.getSpecter()]

    def _rotation_tubes(self):
        return [(self.tube, 0, self.bowtie, self.start_angle, self.fov, self.mas_a),
                (self.tubeB, 1, self.bowtieB or self.bowtie, self.start_angle + self.tube_b_offset, self.fov_b, self.mas_b)]


class CTSequentialBeam(_BeamBase):
    """CTSequentialBeam<false> with the reference's creation defaults (beamsettingsmodel.cpp:925-933)."""

    def __init__(self, position=(0.0, 0.0, 0.0), scan_normal=(0.0, 0.0, 1.0), filtration=None):
        self.position = tuple(position)
        self.scan_normal = tuple(scan_normal)
        self.tube = Tube(voltage=120.0, filtration=dict(filtration if filtration is not None else {13: 9.0}))
        self.slices = 1
        self.slice_spacing = 1.0
        self.start_angle = 0.0
        self.step_angle = math.radians(5.0)
        self.collimation = 3.84
        self.sdd = 119.0
        self.fov = 50.0
        self.bowtie: BowtieFilter | None = None

    def setNumberOfSlices(self, n): self.slices = int(n)
    def setSliceSpacing(self, s): self.slice_spacing = float(s)
    def setStepAngleDeg(self, a): self.step_angle = math.radians(a)
    def setStartAngleDeg(self, a): self.start_angle = math.radians(a)
    def setCollimation(self, c): self.collimation = float(c)
    def setSourceDetectorDistance(self, d): self.sdd = float(d)
    def setScanFieldOfView(self, f): self.fov = float(f)
    def setBowtieFilter(self, b): self.bowtie = b
    def setTubeVoltage(self, kv): self.tube.voltage = float(kv)

    def bowties(self):
        return [self.bowtie] if self.bowtie is not None else []

    def exposures(self) -> np.ndarray:
        p = N.CTSequentialParams(tuple(self.position), tuple(self.scan_normal), self.slices, self.slice_spacing,
                                 self.start_angle, self.step_angle, self.collimation, self.sdd, self.fov,
                                 self.numberOfParticlesPerExposure, int(self.bowtie is not None), 0)
        return self._collect(N.lib().dxg_ct_sequential_exposures, p)

    ctdi_w = 0.0
    ctdi_diameter = 32.0

    def setCTDIw(self, mgy: float) -> None:
        """CTDIw (mGy) of one rotation (beamsettingsmodel.cpp:1050-1060)."""
        self.ctdi_w = float(mgy)

    def CTDIw(self) -> float:
        return self.ctdi_w

    def setCTDIdiameter(self, cm: float) -> None:
        self.ctdi_diameter = float(cm)

    def CTDIdiameter(self) -> float:
        return self.ctdi_diameter

    def _rotation_tubes(self):
        return [(self.tube, 0, self.bowtie, self.start_angle, self.fov, 1.0)]

    def calibrationFactor(self) -> float:
        return _ctdi_factor(self, self.ctdi_w) if self.ctdi_w > 0 else 0.0


CTDI_PARTICLES_PER_EXPOSURE = 100_000


class _Rotation(_BeamBase):
    """One axial rotation at the origin of a CT beam's tubes (CTDI phantom run)."""

    def __init__(self, ct):
        self.ct = ct
        parts, self._spectra, self._bowties = [], [], []
        for tube, idx, bow, start, fov, mas in ct._rotation_tubes():
            p = N.CTSequentialParams((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1, 1.0, start, ct.step_angle, ct.collimation, ct.sdd,
                                     fov, CTDI_PARTICLES_PER_EXPOSURE, int(bow is not None), 0)
            ex = self._collect(N.lib().dxg_ct_sequential_exposures, p)
            ex["spectrum_index"] = len(self._spectra)
            self._spectra.append(tube.getSpecter())
            if bow is not None:
                ex["bowtie_index"] = len(self._bowties)
                self._bowties.append(bow)
            ex["weight"] *= mas
            parts.append(ex)
        self._ex = np.concatenate(parts)

    def exposures(self):
        return self._ex

    def spectra(self):
        return self._spectra

    def bowties(self):
        return self._bowties


def _ctdi_factor(ct, ctdi_w: float) -> float:
    key = (ctdi_w, ct.ctdi_diameter, ct.numberOfParticlesPerExposure, ct.collimation, ct.sdd, ct.step_angle,
           repr([(t.voltage, t.anode_angle_deg, sorted(t.filtration.items()), i, id(b), a, f, m)
                 for t, i, b, a, f, m in ct._rotation_tubes()]))
    cache = getattr(ct, "_ctdi_cache", None)
    if cache and cache[0] == key:
        return cache[1]
    k = _Keep()
    b = beam_struct(_Rotation(ct), k, calibrated=False)
    f, sim = C.c_double(), C.c_double()
    N.check(N.lib().dxg_calibration_ctdi(C.byref(b), float(ctdi_w), float(ct.ctdi_diameter), float(ct.collimation),
                                          int(ct.numberOfParticlesPerExposure), 12345, C.byref(f), C.byref(sim)))
    ct._ctdi_cache = (key, f.value)
    return f.value


class CBCTBeam(_BeamBase):
    """CBCTBeam<false> (beamsettingsmodel.cpp:720-900)."""

    def __init__(self, isocenter=(0.0, 0.0, 0.0), rotation_axis=(0.0, 0.0, 1.0), filtration=None):
        self.isocenter = tuple(isocenter)
        self.rotation_axis = tuple(rotation_axis)
        self.tube = Tube(voltage=90.0, filtration=dict(filtration if filtration is not None else {13: 2.0}))
        self.start_angle = math.radians(-90.0)
        self.stop_angle = math.radians(90.0)
        self.step_angle = math.radians(1.0)
        self.sdd = 120.0
        self.half_angles = (math.radians(10.0), math.radians(10.0))   # setCollimationHalfAnglesDeg(10, 10), :733

    def setStartAngleDeg(self, a): self.start_angle = math.radians(a)
    def setStopAngleDeg(self, a): self.stop_angle = math.radians(a)
    def setStepAngleDeg(self, a): self.step_angle = math.radians(a)
    def setSourceDetectorDistance(self, d): self.sdd = float(d)
    def setCollimationHalfAnglesDeg(self, a, b): self.half_angles = (math.radians(a), math.radians(b))

    dap = 0.0

    def setDAPvalue(self, mgycm2: float) -> None:
        """Dose-area product (mGy cm2) of the whole scan (beamsettingsmodel.cpp:859-868)."""
        self.dap = float(mgycm2)

    def DAPvalue(self) -> float:
        return self.dap

    def calibrationFactor(self) -> float:
        return _calibrate(N.lib().dxg_calibration_dap, self, self.dap) if self.dap > 0 else 0.0

    def exposures(self) -> np.ndarray:
        p = N.CBCTParams(tuple(self.isocenter), tuple(self.rotation_axis), self.start_angle, self.stop_angle,
                         self.step_angle, self.sdd, tuple(self.half_angles), self.numberOfParticlesPerExposure)
        return self._collect(N.lib().dxg_cbct_exposures, p)


# ------------------------------------------------------------------ transport
def beam_struct(beam, keep: _Keep, calibrated: bool = True) -> N.Beam:
    """Flatten a beam object into the dxg_beam record (arrays kept alive by
    `keep`); with `calibrated` the beam's calibration factor is included."""
    ex = keep.keep(np.ascontiguousarray(beam.exposures()))
    specs = []
    for e, w in beam.spectra():
        e = keep.keep(np.ascontiguousarray(e, dtype=np.float64))
        w = keep.keep(np.ascontiguousarray(w, dtype=np.float64))
        specs.append(N.Spectrum(N.ptr(e, C.c_double), N.ptr(w, C.c_double), len(e)))
    spec_arr = keep.keep((N.Spectrum * len(specs))(*specs))
    bows = []
    for b in beam.bowties():
        a = keep.keep(np.ascontiguousarray(b.angle, dtype=np.float64))
        w = keep.keep(np.ascontiguousarray(b.weight, dtype=np.float64))
        bows.append(N.Bowtie(N.ptr(a, C.c_double), N.ptr(w, C.c_double), len(a)))
    bow_arr = keep.keep((N.Bowtie * max(1, len(bows)))(*bows))
    cal = beam.calibrationFactor() if calibrated else 0.0
    return N.Beam(ex.ctypes.data_as(C.POINTER(N.Exposure)), len(ex), spec_arr, len(specs), bow_arr, len(bows), cal)


class Transport:
    """dxmc::Transport: blocking run of one beam into a world."""

    def __init__(self, seed: int = 1):
        self.seed = int(seed)
        self.threads = 0

    def setNumberOfThreads(self, n: int) -> None:
        self.threads = int(n)   # host threads do not drive the device path

    def __call__(self, world: World, beam, progress: TransportProgress | None = None, useBeamCalibration: bool = True):
        k = _Keep()
        b = beam_struct(beam, k, calibrated=useBeamCalibration)
        grid = world.items[0]
        rc = N.lib().dxg_transport_run(grid._h, C.byref(b), progress._h if progress else None,
                                       int(bool(useBeamCalibration)), self.seed)
        if rc == N.DXG_ERR_CANCELLED:
            return False
        N.check(rc)
        return True


def sample_source(world: World, beam, n: int, seed: int = 1) -> np.ndarray:
    """Source particles of beam histories 0..n-1 from the transport kernels'
    own sampling code: an (n, 8) float32 array of ax, ay, direction (3),
    energy (keV), weight and exposure index."""
    k = _Keep()
    b = beam_struct(beam, k, calibrated=False)
    out = np.zeros((int(n), 8), dtype=np.float32)
    N.check(N.lib().dxg_debug_source(world.items[0]._h, C.byref(b), int(seed), int(n), N.ptr(out, C.c_float)))
    return out


def compton_cos(energy_kev: float, u, material: dict[int, float] | None = None) -> np.ndarray:
    """cos(theta) of Compton scatters at energy_kev (1-200 keV) for uniforms u
    in [0, 1), by the transport kernels' tables and FP32 interpolation,
    evaluated on the host (no device needed): Klein-Nishina (correction
    level 0), or with a material composition Klein-Nishina x S(x)/Z of that
    material (levels 1 and 2)."""
    u = np.ascontiguousarray(u, dtype=np.float32).ravel()
    out = np.zeros_like(u)
    if material is None:
        N.check(N.lib().dxg_debug_compton_cos(float(energy_kev), u.size, N.ptr(u, C.c_float), N.ptr(out, C.c_float)))
    else:
        k = _Keep()
        m = Material(material)._struct(k)
        N.check(N.lib().dxg_debug_compton_cos_material(C.byref(m), float(energy_kev), u.size, N.ptr(u, C.c_float),
                                                       N.ptr(out, C.c_float)))
    return out


class TransportBatch:
    """Many (world, beam) instances in one persistent device launch, replayed
    from a captured HIP graph (no counterpart in the reference).  Each
    instance's scores are read from its grid afterwards."""

    def __init__(self):
        h = C.c_void_p()
        N.check(N.lib().dxg_batch_create(C.byref(h)))
        self._h = h
        self._worlds = []

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().dxg_batch_destroy(self._h)
            self._h = None

    def add(self, world: World, beam, seed: int = 1) -> None:
        k = _Keep()
        b = beam_struct(beam, k)
        N.check(N.lib().dxg_batch_add(self._h, world.items[0]._h, C.byref(b), int(seed)))
        self._worlds.append(world)      # the worlds must outlive the batch

    def __len__(self) -> int:
        return int(N.lib().dxg_batch_size(self._h))

    def run(self) -> float:
        """One launch over every instance; returns its device time in ms."""
        ms = C.c_double()
        N.check(N.lib().dxg_batch_run(self._h, C.byref(ms)))
        return ms.value

