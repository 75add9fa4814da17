"""ctypes binding of libdxmc_hip.so (include/dxmc_hip.h).

The library is the only compute path: if it is missing, importing the binding
raises, and every entry point reports device errors as exceptions.  There is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# DXG_LIBRARY points at another build of the library (same C-ABI)
LIB_PATH = Path(os.environ.get("DXG_LIBRARY") or Path(__file__).resolve().parent / "libdxmc_hip.so")

DXG_OK = 0
DXG_ERR_ARGUMENT = -1
DXG_ERR_STATE = -2
DXG_ERR_MATERIAL = -3
DXG_ERR_DATA = -4
DXG_ERR_PHYSICS = -5
DXG_ERR_DEVICE = -6
DXG_ERR_UNSUPPORTED = -7
DXG_ERR_CANCELLED = -8

u64 = C.c_uint64
i64 = C.c_int64
PD = C.POINTER(C.c_double)


class Material(C.Structure):
    _fields_ = [("Z", C.POINTER(C.c_int32)), ("mass_fraction", PD), ("n", u64)]


class Exposure(C.Structure):
    _fields_ = [
        ("position", C.c_double * 3),
        ("cosines", C.c_double * 6),
        ("collimation_half_angles", C.c_double * 2),
        ("weight", C.c_double),
        ("spectrum_index", C.c_uint32),
        ("bowtie_index", C.c_int32),
        ("n_particles", u64),
    ]


class Spectrum(C.Structure):
    _fields_ = [("energy_kev", PD), ("weight", PD), ("n", u64)]


class Bowtie(C.Structure):
    _fields_ = [("angle_rad", PD), ("weight", PD), ("n", u64)]


class Beam(C.Structure):
    _fields_ = [
        ("exposures", C.POINTER(Exposure)), ("n_exposures", u64),
        ("spectra", C.POINTER(Spectrum)), ("n_spectra", u64),
        ("bowties", C.POINTER(Bowtie)), ("n_bowties", u64),
        ("calibration", C.c_double),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("histories", u64), ("steps", u64), ("real_events", u64), ("deposits", u64),
        ("emitted_energy", C.c_double), ("escaped_energy", C.c_double),
        ("deposited_energy", C.c_double), ("kernel_ms", C.c_double), ("kernel_launches", u64),
        ("uncollided", u64), ("cell_crossings", u64), ("phase_cycles", u64 * 4),
        ("phase_lanes", u64 * 4), ("phase_iters", u64 * 4),
        ("refill_threshold", C.c_int32), ("interact_threshold", C.c_int32), ("outer_threshold", C.c_int32),
        ("autotuned", C.c_int32), ("batch", u64),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        for name in ("phase_cycles", "phase_lanes", "phase_iters"):
            d[name] = list(getattr(self, name))
        return d


class Tube(C.Structure):
    _fields_ = [("voltage_kv", C.c_double), ("anode_angle_deg", C.c_double), ("filtration_Z", C.POINTER(C.c_int32)),
                ("filtration_mm", PD), ("n_filtration", u64), ("energy_step_kev", C.c_double)]


class AEC(C.Structure):
    _fields_ = [("start", C.c_double * 3), ("stop", C.c_double * 3), ("weights", PD), ("n", u64)]


class OrganAEC(C.Structure):
    _fields_ = [("use", C.c_int32), ("compensate_outside", C.c_int32), ("start_angle", C.c_double),
                ("stop_angle", C.c_double), ("ramp_angle", C.c_double), ("low_weight", C.c_double)]


class PencilParams(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("direction", C.c_double * 3), ("n_exposures", u64),
                ("particles_per_exposure", u64)]


class DXParams(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("cosines", C.c_double * 6), ("collimation_half_angles", C.c_double * 2),
                ("n_exposures", u64), ("particles_per_exposure", u64)]


class CTSpiralParams(C.Structure):
    _fields_ = [("start", C.c_double * 3), ("stop", C.c_double * 3), ("start_angle", C.c_double),
                ("step_angle", C.c_double), ("pitch", C.c_double), ("collimation_cm", C.c_double),
                ("sdd_cm", C.c_double), ("fov_cm", C.c_double), ("particles_per_exposure", u64),
                ("dual_source", C.c_int32), ("has_bowtie", C.c_int32), ("fov_b_cm", C.c_double),
                ("tube_b_offset_angle", C.c_double), ("relative_mas_a", C.c_double), ("relative_mas_b", C.c_double),
                ("aec", AEC), ("organ_aec", OrganAEC)]


class CTSequentialParams(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("scan_normal", C.c_double * 3), ("number_of_slices", u64),
                ("slice_spacing_cm", C.c_double), ("start_angle", C.c_double), ("step_angle", C.c_double),
                ("collimation_cm", C.c_double), ("sdd_cm", C.c_double), ("fov_cm", C.c_double),
                ("particles_per_exposure", u64), ("has_bowtie", C.c_int32), ("pad", C.c_int32)]


class CBCTParams(C.Structure):
    _fields_ = [("isocenter", C.c_double * 3), ("rotation_axis", C.c_double * 3), ("start_angle", C.c_double),
                ("stop_angle", C.c_double), ("step_angle", C.c_double), ("sdd_cm", C.c_double),
                ("collimation_half_angles", C.c_double * 2), ("particles_per_exposure", u64)]


EXPOSURE_DTYPE = np.dtype([
    ("position", "<f8", 3), ("cosines", "<f8", 6), ("collimation_half_angles", "<f8", 2),
    ("weight", "<f8"), ("spectrum_index", "<u4"), ("bowtie_index", "<i4"), ("n_particles", "<u8"),
], align=True)
assert EXPOSURE_DTYPE.itemsize == C.sizeof(Exposure) == 112


SEG_MAX_MATERIALS = 128


class CTSegmentation(C.Structure):
    """dxg_ct_segmentation (include/dxmc_pipelines.h)."""
    _fields_ = [("n_materials", C.c_uint32), ("reserved", C.c_uint32),
                ("hu", C.c_double * SEG_MAX_MATERIALS), ("separators", C.c_double * (SEG_MAX_MATERIALS - 1)),
                ("attenuation", C.c_double * SEG_MAX_MATERIALS),
                ("water_attenuation", C.c_double), ("air_attenuation", C.c_double),
                ("water_density", C.c_double), ("air_density", C.c_double),
                ("slope", C.c_double), ("offset", C.c_double)]


class OrganDose(C.Structure):
    """dxg_organ_dose (include/dxmc_pipelines.h)."""
    _fields_ = [("n_voxels", u64), ("volume_cm3", C.c_double), ("mass_g", C.c_double), ("dose", C.c_double)]


ORGAN_DOSE_DTYPE = np.dtype([("n_voxels", np.uint64), ("volume_cm3", np.float64), ("mass_g", np.float64),
                             ("dose", np.float64)])
assert ORGAN_DOSE_DTYPE.itemsize == C.sizeof(OrganDose) == 32


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

PF = C.POINTER(C.c_float)

# (name, restype, argtypes) for every entry point of include/dxmc_hip.h,
# include/dxmc_beams.h, include/dxmc_denoise.h, include/dxmc_pipelines.h and include/dxmc_h5.h
SIGNATURES = [
    ("dxg_last_error", C.c_char_p, []),
    ("dxg_world_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("dxg_world_destroy", None, [C.c_void_p]),
    ("dxg_grid_set_data", C.c_int, [C.c_void_p, C.POINTER(u64), PD, C.POINTER(C.c_uint8), C.POINTER(Material), u64]),
    ("dxg_grid_set_spacing", C.c_int, [C.c_void_p, PD]),
    ("dxg_world_build", C.c_int, [C.c_void_p]),
    ("dxg_grid_size", u64, [C.c_void_p]),
    ("dxg_transport_run", C.c_int, [C.c_void_p, C.POINTER(Beam), C.c_void_p, C.c_int, u64]),
    ("dxg_transport_run_static", C.c_int, [C.c_void_p, C.POINTER(Beam), C.c_void_p, C.c_int, u64]),
    ("dxg_world_clear_scores", C.c_int, [C.c_void_p]),
    ("dxg_read_energy_scores", C.c_int, [C.c_void_p, PD, PD, C.POINTER(u64)]),
    ("dxg_read_dose_scores", C.c_int, [C.c_void_p, PD, PD, C.POINTER(u64)]),
    ("dxg_world_histories", u64, [C.c_void_p]),
    ("dxg_get_counters", C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    ("dxg_progress_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("dxg_progress_destroy", None, [C.c_void_p]),
    ("dxg_progress_request_stop", None, [C.c_void_p]),
    ("dxg_progress_continue", C.c_int, [C.c_void_p]),
    ("dxg_progress_get", None, [C.c_void_p, C.POINTER(u64), C.POINTER(u64)]),
    ("dxg_progress_message", None, [C.c_void_p, C.c_char_p, u64]),
    ("dxg_material_attenuation", C.c_int, [C.POINTER(Material), PD, u64, PD]),
    ("dxg_tube_spectrum", C.c_int, [C.POINTER(Tube), PD, PD, u64, C.POINTER(u64)]),
    ("dxg_tube_hvl_mm_al", C.c_int, [C.POINTER(Tube), PD]),
    ("dxg_tube_mean_energy", C.c_int, [C.POINTER(Tube), PD]),
    ("dxg_organ_aec_weight", C.c_int, [C.POINTER(OrganAEC), C.c_double, PD]),
    ("dxg_organ_aec_max_weight", C.c_int, [C.POINTER(OrganAEC), PD]),
    ("dxg_aec_from_wed", C.c_int, [C.POINTER(u64), PD, PD, C.c_int, PD, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("dxg_pencil_exposures", C.c_int, [C.POINTER(PencilParams), C.POINTER(Exposure), u64, C.POINTER(u64)]),
    ("dxg_dx_exposures", C.c_int, [C.POINTER(DXParams), C.POINTER(Exposure), u64, C.POINTER(u64)]),
    ("dxg_ct_spiral_exposures", C.c_int, [C.POINTER(CTSpiralParams), C.POINTER(Exposure), u64, C.POINTER(u64)]),
    ("dxg_ct_sequential_exposures", C.c_int, [C.POINTER(CTSequentialParams), C.POINTER(Exposure), u64, C.POINTER(u64)]),
    ("dxg_cbct_exposures", C.c_int, [C.POINTER(CBCTParams), C.POINTER(Exposure), u64, C.POINTER(u64)]),
    ("dxg_debug_device_copy", C.c_int, [u64, C.c_int, PD]),
    ("dxg_debug_traverse", C.c_int, [C.c_void_p, u64, C.POINTER(C.c_float), C.POINTER(C.c_float), u64,
                                     C.POINTER(C.c_float), C.POINTER(i64), C.POINTER(C.c_int32)]),
    ("dxg_debug_source", C.c_int, [C.c_void_p, C.POINTER(Beam), u64, u64, C.POINTER(C.c_float)]),
    ("dxg_debug_compton_cos", C.c_int, [C.c_double, u64, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("dxg_debug_compton_cos_material", C.c_int, [C.POINTER(Material), C.c_double, u64, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("dxg_calibration_dap", C.c_int, [C.POINTER(Beam), C.c_double, PD]),
    ("dxg_calibration_air_kerma", C.c_int, [C.POINTER(Beam), C.c_double, PD]),
    ("dxg_calibration_ctdi", C.c_int, [C.POINTER(Beam), C.c_double, C.c_double, C.c_double, u64, u64, PD, PD]),
    ("dxg_batch_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("dxg_batch_destroy", None, [C.c_void_p]),
    ("dxg_batch_add", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Beam), u64]),
    ("dxg_batch_size", u64, [C.c_void_p]),
    ("dxg_batch_run", C.c_int, [C.c_void_p, PD]),
    ("dxg_denoiser_num_params", u64, []),
    ("dxg_denoiser_create", C.c_int, [C.POINTER(C.c_void_p), u64]),
    ("dxg_denoiser_destroy", None, [C.c_void_p]),
    ("dxg_denoiser_get_params", C.c_int, [C.c_void_p, PF, u64]),
    ("dxg_denoiser_set_params", C.c_int, [C.c_void_p, PF, u64]),
    ("dxg_denoiser_get_grads", C.c_int, [C.c_void_p, PF, u64]),
    ("dxg_denoiser_train_step", C.c_int, [C.c_void_p, PF, PF, PF, u64, C.POINTER(u64), C.c_float, PF]),
    ("dxg_denoiser_infer", C.c_int, [C.c_void_p, PF, C.POINTER(u64), PF]),
    ("dxg_denoiser_last_ms", C.c_double, [C.c_void_p]),
    ("dxg_denoiser_world_features", C.c_int, [C.c_void_p, C.c_void_p, PF, u64]),
    ("dxg_denoiser_apply", C.c_int, [C.c_void_p, C.c_void_p, PD]),
    ("dxg_denoiser_debug_tiles", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int32), u64, C.POINTER(u64)]),
    ("dxg_conv3d_probe", C.c_int, [C.c_int, PF, PF, PF, u64, C.POINTER(u64), PF]),
    ("dxg_mfma_probe", C.c_int, [PF, PF, PF]),
    ("dxg_ct_segmentation_tables", C.c_int, [C.POINTER(Tube), C.POINTER(Material), PD, u64, C.POINTER(Material),
                                             C.c_double, C.POINTER(Material), C.c_double, C.POINTER(CTSegmentation)]),
    ("dxg_ct_segment", C.c_int, [C.POINTER(CTSegmentation), PD, u64, C.POINTER(C.c_uint8), PD]),
    ("dxg_ct_segment_device", C.c_int, [C.POINTER(CTSegmentation), C.c_void_p, u64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("dxg_organ_dose_table", C.c_int, [C.POINTER(C.c_uint8), PD, PD, u64, PD, C.POINTER(OrganDose)]),
    ("dxg_organ_dose_table_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, u64, PD, C.POINTER(OrganDose),
                                              C.c_void_p]),
    ("dxg_pipeline_last_ms", C.c_double, [C.c_int]),
    ("dxg_h5_create", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("dxg_h5_open", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("dxg_h5_close", C.c_int, [C.c_void_p]),
    ("dxg_h5_create_group", C.c_int, [C.c_void_p, C.c_char_p]),
    ("dxg_h5_kind", C.c_int, [C.c_void_p, C.c_char_p]),
    ("dxg_h5_write_array", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(u64), C.c_void_p, C.c_int]),
    ("dxg_h5_write_strings", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_char_p), u64]),
    ("dxg_h5_write_attribute", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, u64, C.c_int]),
    ("dxg_h5_dataset_shape", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(C.c_int),
                                       C.POINTER(u64)]),
    ("dxg_h5_read_array", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p, u64]),
    ("dxg_h5_read_strings", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, u64, C.POINTER(u64), C.POINTER(u64)]),
    ("dxg_h5_read_attribute", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, u64, C.POINTER(u64),
                                        C.POINTER(C.c_int)]),
    ("dxg_h5_list", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, u64, C.POINTER(u64), C.POINTER(u64)]),
    ("dxg_h5_last_deflate_ms", C.c_double, [C.POINTER(C.c_int)]),
]


def lib() -> C.CDLL:
    """Load libdxmc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        handle = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != DXG_OK:
        raise NativeError(rc, (lib().dxg_last_error() or b"").decode())


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
