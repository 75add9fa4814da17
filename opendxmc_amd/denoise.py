"""Dose denoiser: a 3-D residual convolutional network that maps a low-history
dose map to a high-history one (include/dxmc_denoise.h; BF16 matrix-core
kernels in libdxmc_hip.so, no CPU path).

The reference has no counterpart; its plain Monte Carlo result is what the
network approximates, and a high-history run of this engine is both its
training target and its fallback.

Features per voxel (channel order of the network input):
  0  relative dose          D / D_ref, D_ref = mean dose over the body (material != 0)
  1  density (g/cm3)
  2  material embedding     mass attenuation at 60 keV relative to water
  3  relative standard error of the dose (1 where no event was scored)
The network adds a correction to feature 0 and so predicts the relative dose
of the high-history run; the dose follows as max(pred, 0) * D_ref.  Working on
the linear dose keeps the squared-error optimum at the conditional mean, so a
Monte Carlo target with noise of its own still trains an unbiased network.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as N

FEATURES = 4
LAYERS = 18
CHANNELS = 32
TAPS = 27
LAYER_W = CHANNELS * TAPS * CHANNELS
LAYER_P = LAYER_W + CHANNELS

_WATER = {1: 0.111894, 8: 0.888106}

# Network trained by denoise_study on engine-generated CT spiral pairs at
# 0.08 cm voxels (12 seeded body phantoms of 448 x 448 x 192, inputs at 1-100
# histories per voxel, targets at 400; 5000 Adam steps on 64^3 patches).
PRETRAINED_CT = Path(__file__).resolve().parent / "data" / "denoiser_ct_0p08cm.npy"


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Denoiser:
    """Owns the network's device state (FP32 master weights, Adam moments)."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        N.check(N.lib().dxg_denoiser_create(C.byref(h), int(seed)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().dxg_denoiser_destroy(h)
            self._h = None

    @staticmethod
    def num_params() -> int:
        return int(N.lib().dxg_denoiser_num_params())

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params(), dtype=np.float32)
        N.check(N.lib().dxg_denoiser_get_params(self._h, _pf(p), p.size))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().dxg_denoiser_set_params(self._h, _pf(p), p.size))

    def grads(self) -> np.ndarray:
        g = np.zeros(self.num_params(), dtype=np.float32)
        N.check(N.lib().dxg_denoiser_get_grads(self._h, _pf(g), g.size))
        return g

    def train_step(self, features: np.ndarray, target: np.ndarray, mask: np.ndarray, lr: float = 2e-4) -> float:
        """One Adam step on a batch of patches: features (B, D, H, W, 4),
        target and mask (B, D, H, W); W a multiple of 64, H of 4."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        if f.ndim != 5 or f.shape[-1] != FEATURES:
            raise ValueError("features must be (batch, D, H, W, 4)")
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        if t.shape != f.shape[:-1] or m.shape != f.shape[:-1]:
            raise ValueError("target and mask must be (batch, D, H, W)")
        dims = np.array(f.shape[1:4], dtype=np.uint64)
        loss = C.c_float()
        N.check(N.lib().dxg_denoiser_train_step(self._h, _pf(f), _pf(t), _pf(m), f.shape[0],
                                                 dims.ctypes.data_as(C.POINTER(N.u64)), float(lr), C.byref(loss)))
        return float(loss.value)

    def infer(self, features: np.ndarray) -> np.ndarray:
        """Predicted relative dose for one volume (D, H, W, 4); W a multiple
        of 64 (see pad_x)."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        if f.ndim != 4 or f.shape[-1] != FEATURES:
            raise ValueError("features must be (D, H, W, 4)")
        out = np.zeros(f.shape[:3], dtype=np.float32)
        dims = np.array(f.shape[:3], dtype=np.uint64)
        N.check(N.lib().dxg_denoiser_infer(self._h, _pf(f), dims.ctypes.data_as(C.POINTER(N.u64)), _pf(out)))
        return out

    def last_ms(self) -> float:
        """Device time of the last train_step / infer / apply call (HIP events)."""
        return float(N.lib().dxg_denoiser_last_ms(self._h))

    def world_features(self, grid) -> np.ndarray:
        """Features of a grid's current scores computed on the device,
        (nz, ny, nx padded to 64, 4)."""
        nx, ny, nz = grid.dims
        nxp = (nx + 63) // 64 * 64
        f = np.zeros((nz, ny, nxp, FEATURES), dtype=np.float32)
        N.check(N.lib().dxg_denoiser_world_features(self._h, grid._h, _pf(f), f.size))
        return f

    def apply(self, grid, copy_out: bool = True) -> np.ndarray | None:
        """Denoise a grid's current dose on the device; returns the dose (mGy,
        flat grid order) unless copy_out is False (device timing only)."""
        out = np.zeros(grid.size(), dtype=np.float64) if copy_out else None
        p = out.ctypes.data_as(C.POINTER(C.c_double)) if copy_out else None
        N.check(N.lib().dxg_denoiser_apply(self._h, grid._h, p))
        return out

    def debug_tiles(self, grid, host: bool = False, layer: int = 0) -> np.ndarray:
        """The body-tile list apply() runs convolution `layer` on (row-pair
        entries pair << 2 | rows to store; the tiles within 19 - layer voxels
        of the body), built on the device or, with host=True, by the host scan."""
        n = N.u64(0)
        N.check(N.lib().dxg_denoiser_debug_tiles(self._h, grid._h, int(host), int(layer), None, 0, C.byref(n)))
        out = np.zeros(max(int(n.value), 1), dtype=np.int32)
        N.check(N.lib().dxg_denoiser_debug_tiles(self._h, grid._h, int(host), int(layer),
                                                 out.ctypes.data_as(C.POINTER(C.c_int32)), int(n.value), C.byref(n)))
        return out[:int(n.value)]

    @classmethod
    def pretrained(cls, path=PRETRAINED_CT) -> "Denoiser":
        net = cls(seed=1)
        net.load(path)
        return net

    def save(self, path) -> None:
        np.save(path, self.params())

    def load(self, path) -> None:
        self.set_params(np.load(path, allow_pickle=False))


def material_embedding(materials: list[dict[int, float]], energy_kev: float = 60.0) -> np.ndarray:
    """Per-material scalar: total mass attenuation at `energy_kev` over water's."""
    from .engine import Material

    water = Material(_WATER).attenuationValues(np.array([energy_kev]))[0, 3]
    return np.array([Material(m).attenuationValues(np.array([energy_kev]))[0, 3] / water for m in materials],
                     dtype=np.float32)


@dataclass
class DoseVolume:
    """One Monte Carlo result on a grid, arrays shaped (nz, ny, nx)."""
    dose: np.ndarray           # mGy (or any unit shared by the pair)
    variance: np.ndarray       # variance of the dose, same unit squared
    density: np.ndarray        # g/cm3
    material: np.ndarray       # material index
    embedding: np.ndarray      # per-material scalar (material_embedding)

    @staticmethod
    def from_grid(grid, dims, density, material, embedding) -> "DoseVolume":
        nx, ny, nz = (int(v) for v in dims)
        d, v, _ = grid.doseScored()
        shape = (nz, ny, nx)
        return DoseVolume(d.reshape(shape), v.reshape(shape), np.asarray(density).reshape(shape),
                          np.asarray(material).reshape(shape), np.asarray(embedding, dtype=np.float32))


def dose_reference(dose: np.ndarray, material: np.ndarray) -> float:
    body = material != 0
    ref = float(dose[body].mean()) if body.any() else float(dose.mean())
    return ref if ref > 0 else 1.0


def features(vol: DoseVolume, dose_ref: float | None = None) -> tuple[np.ndarray, float]:
    """(nz, ny, nx, 4) float32 network input and the dose normalisation."""
    if dose_ref is None:
        dose_ref = dose_reference(vol.dose, vol.material)
    d = np.maximum(vol.dose, 0.0)
    f = np.empty(vol.dose.shape + (FEATURES,), dtype=np.float32)
    f[..., 0] = d / dose_ref
    f[..., 1] = vol.density
    f[..., 2] = vol.embedding[vol.material.astype(np.intp)]
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(d > 0, np.sqrt(np.maximum(vol.variance, 0.0)) / d, 1.0)
    f[..., 3] = np.clip(rel, 0.0, 1.0)
    return f, dose_ref


def relative_dose(dose: np.ndarray, dose_ref: float) -> np.ndarray:
    return (np.maximum(dose, 0.0) / dose_ref).astype(np.float32)


def dose_from_relative(pred: np.ndarray, dose_ref: float) -> np.ndarray:
    return np.maximum(pred.astype(np.float64), 0.0) * dose_ref


def pad_x(a: np.ndarray, multiple: int = 64, axis: int = 2) -> np.ndarray:
    """Zero-pad the x axis (axis 2 of (nz, ny, nx[, c])) to a multiple."""
    n = a.shape[axis]
    extra = (-n) % multiple
    if extra == 0:
        return a
    pad = [(0, 0)] * a.ndim
    pad[axis] = (0, extra)
    return np.pad(a, pad)


def denoise(net: Denoiser, vol: DoseVolume) -> np.ndarray:
    """Denoised dose (nz, ny, nx) of one low-history result."""
    f, ref = features(vol)
    nx = f.shape[2]
    pred = net.infer(pad_x(f))[:, :, :nx]
    return dose_from_relative(pred, ref)


def mean_relative_error(dose: np.ndarray, reference: np.ndarray, threshold: float = 0.2) -> float:
    """Mean |dose - reference| / reference over voxels above threshold x max(reference)."""
    sel = reference > threshold * reference.max()
    return float(np.mean(np.abs(dose[sel] - reference[sel]) / reference[sel]))
