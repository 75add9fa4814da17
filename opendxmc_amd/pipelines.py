"""CT segmentation and organ dose table on the device (include/dxmc_pipelines.h)
behind OpenDXMC's pipeline names: CTSegmentationPipeline
(ctsegmentationpipeline.cpp) and DoseTablePipeline (dosetablepipeline.cpp).

Host arrays (numpy) go through the staging entry points; device-resident
volumes (addresses from opendxmc_amd.devmem or any allocation of the library's
HIP runtime) go through the _device entry points and stay on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .engine import Material, Tube, _Keep

# NIST compositions (mass fractions) and densities (g/cm3) of the materials
# CTSegmentationPipeline names (ctsegmentationpipeline.cpp:120), restated from
# NIST's published material composition tables under DXMClib's names.
NIST_MATERIALS: dict[str, tuple[float, dict[int, float]]] = {
    "Air, Dry (near sea level)": (1.20479e-3, {6: 0.000124, 7: 0.755268, 8: 0.231781, 18: 0.012827}),
    "Water, Liquid": (1.0, {1: 0.111894, 8: 0.888106}),
    "Adipose Tissue (ICRP)": (0.92, {1: 0.119477, 6: 0.637240, 7: 0.007970, 8: 0.232333, 11: 0.000500, 12: 0.000020,
                                     15: 0.000160, 16: 0.000730, 17: 0.001190, 19: 0.000320, 20: 0.000020,
                                     26: 0.000020, 30: 0.000020}),
    "Tissue, Soft (ICRP)": (1.0, {1: 0.104472, 6: 0.232190, 7: 0.024880, 8: 0.630238, 11: 0.001130, 12: 0.000130,
                                  15: 0.001330, 16: 0.001990, 17: 0.001340, 19: 0.001990, 20: 0.000230,
                                  26: 0.000050, 30: 0.000030}),
    "Muscle, Skeletal": (1.04, {1: 0.100637, 6: 0.107830, 7: 0.027680, 8: 0.754773, 11: 0.000750, 12: 0.000190,
                                15: 0.001800, 16: 0.002410, 17: 0.000790, 19: 0.003020, 20: 0.000030,
                                26: 0.000040, 30: 0.000050}),
    "Bone, Cortical (ICRP)": (1.85, {1: 0.047234, 6: 0.144330, 7: 0.041990, 8: 0.446096, 12: 0.002200,
                                     15: 0.104970, 16: 0.003150, 20: 0.209930, 30: 0.000100}),
}

# ctsegmentationpipeline.cpp:120-127: the candidate materials in separator
# order; cortical bone at 1.09 g/cm3 instead of its NIST density.
SEGMENTATION_MATERIALS = ["Air, Dry (near sea level)", "Adipose Tissue (ICRP)", "Tissue, Soft (ICRP)",
                          "Muscle, Skeletal", "Bone, Cortical (ICRP)"]
SEGMENTATION_DENSITY_OVERRIDE = {"Bone, Cortical (ICRP)": 1.09}


def segmentation_tables(tube: Tube, materials: list[dict[int, float]], densities: list[float],
                        water: tuple[float, dict[int, float]] | None = None,
                        air: tuple[float, dict[int, float]] | None = None) -> N.CTSegmentation:
    """dxg_ct_segmentation_tables: CT numbers, separators and mean attenuation
    of `materials` for `tube`'s spectrum; water and air default to NIST's."""
    water = water or NIST_MATERIALS["Water, Liquid"]
    air = air or NIST_MATERIALS["Air, Dry (near sea level)"]
    k = _Keep()
    mats = (N.Material * len(materials))(*[Material(m)._struct(k) for m in materials])
    dens = k.keep(np.ascontiguousarray(densities, dtype=np.float64))
    w = Material(water[1])._struct(k)
    a = Material(air[1])._struct(k)
    t = tube._struct(k)
    out = N.CTSegmentation()
    N.check(N.lib().dxg_ct_segmentation_tables(C.byref(t), mats, N.ptr(dens, C.c_double), len(materials), C.byref(w),
                                               float(water[0]), C.byref(a), float(air[0]), C.byref(out)))
    return out


def segment(tables: N.CTSegmentation, ct) -> tuple[np.ndarray, np.ndarray]:
    """dxg_ct_segment: (material uint8, density float64) of the same shape as
    the CT numbers `ct` (host array)."""
    hu = np.ascontiguousarray(ct, dtype=np.float64)
    mat = np.empty(hu.shape, dtype=np.uint8)
    dens = np.empty(hu.shape, dtype=np.float64)
    N.check(N.lib().dxg_ct_segment(C.byref(tables), N.ptr(hu, C.c_double), hu.size, N.ptr(mat, C.c_uint8),
                                   N.ptr(dens, C.c_double)))
    return mat, dens


def segment_device(tables: N.CTSegmentation, hu: int, n: int, material: int, density: int, stream: int | None = None) -> None:
    """dxg_ct_segment_device on device addresses (float64 CT numbers in,
    uint8 materials and float64 densities out)."""
    N.check(N.lib().dxg_ct_segment_device(C.byref(tables), C.c_void_p(hu), int(n), C.c_void_p(material),
                                          C.c_void_p(density), C.c_void_p(stream)))


def _spacing(spacing) -> np.ndarray:
    sp = np.ascontiguousarray(spacing, dtype=np.float64)
    if sp.shape != (3,):
        raise ValueError("spacing must hold three values (cm)")
    return sp


def _rows(rows) -> np.ndarray:
    return np.frombuffer(bytes(rows), dtype=N.ORGAN_DOSE_DTYPE).copy()


def organ_dose_table(organ, dose, density, spacing) -> np.ndarray:
    """dxg_organ_dose_table: 256 rows (n_voxels, volume_cm3, mass_g, dose)
    indexed by organ, from host arrays."""
    sp = _spacing(spacing)
    o = np.ascontiguousarray(organ, dtype=np.uint8).ravel()
    d = np.ascontiguousarray(dose, dtype=np.float64).ravel()
    r = np.ascontiguousarray(density, dtype=np.float64).ravel()
    if not (o.size == d.size == r.size):
        raise ValueError("organ, dose and density must have the same number of voxels")
    rows = (N.OrganDose * 256)()
    N.check(N.lib().dxg_organ_dose_table(N.ptr(o, C.c_uint8), N.ptr(d, C.c_double), N.ptr(r, C.c_double), o.size,
                                         N.ptr(sp, C.c_double), rows))
    return _rows(rows)


def organ_dose_table_device(organ: int, dose: int, density: int, n: int, spacing, stream: int | None = None) -> np.ndarray:
    """dxg_organ_dose_table_device on device addresses (uint8 organ, float64
    dose and density)."""
    sp = _spacing(spacing)
    rows = (N.OrganDose * 256)()
    N.check(N.lib().dxg_organ_dose_table_device(C.c_void_p(organ), C.c_void_p(dose), C.c_void_p(density), int(n),
                                                N.ptr(sp, C.c_double), rows, C.c_void_p(stream)))
    return _rows(rows)


def last_ms(which: int) -> float:
    """Duration of the last segmentation (0) or dose table (1) launch on this thread."""
    return float(N.lib().dxg_pipeline_last_ms(which))


@dataclass
class Segmentation:
    material: object             # uint8 per voxel
    density: object              # g/cm3 per voxel
    names: list[str]             # DataContainer::Material::name
    compositions: list[dict[int, float]]   # DataContainer::Material::Z (mass fractions)


class CTSegmentationPipeline:
    """CTSegmentationPipeline (ctsegmentationpipeline.hpp): acquisition kV,
    Al and Sn filtration (mm) of the CT series; updateImageData maps its CT
    numbers to the five materials and their densities."""

    def __init__(self):
        self.kv = 120.0
        self.al_mm = 0.9
        self.sn_mm = 0.0

    def setAqusitionVoltage(self, kv: float) -> None:
        self.kv = float(kv)

    def setAlFiltration(self, mm: float) -> None:
        self.al_mm = max(0.0, float(mm))

    def setSnFiltration(self, mm: float) -> None:
        self.sn_mm = max(0.0, float(mm))

    def tube(self) -> Tube:
        return Tube(self.kv, filtration={13: self.al_mm, 50: self.sn_mm})

    def tables(self) -> N.CTSegmentation:
        mats = [NIST_MATERIALS[n][1] for n in SEGMENTATION_MATERIALS]
        dens = [SEGMENTATION_DENSITY_OVERRIDE.get(n, NIST_MATERIALS[n][0]) for n in SEGMENTATION_MATERIALS]
        return segmentation_tables(self.tube(), mats, dens)

    def updateImageData(self, ct) -> Segmentation | None:
        if ct is None:
            return None
        mat, dens = segment(self.tables(), ct)
        return Segmentation(mat, dens, list(SEGMENTATION_MATERIALS),
                            [dict(NIST_MATERIALS[n][1]) for n in SEGMENTATION_MATERIALS])


class DoseTablePipeline:
    """DoseTablePipeline (dosetablepipeline.cpp:34-87): one row per named
    organ that has voxels."""

    @staticmethod
    def header(units: str = "mGy") -> list[str]:
        return ["Name", "# Voxels", "Volume cm3", "Mass g", "Dose " + units]

    def updateImageData(self, organ, organ_names: list[str], dose, density, spacing) -> list[tuple]:
        if organ is None or dose is None:
            return []
        table = organ_dose_table(organ, dose, density, spacing)
        rows = []
        for i, name in enumerate(organ_names[:256]):
            r = table[i]
            if r["n_voxels"] > 0:
                rows.append((name, int(r["n_voxels"]), float(r["volume_cm3"]), float(r["mass_g"]), float(r["dose"])))
        return rows
