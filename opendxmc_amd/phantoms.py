"""Synthetic voxel phantoms for the benchmark and test workloads.

Material compositions are the reference's own ICRP adult-male media
(data/phantoms/icrp/AM/AM_media.dat, % by mass of H C N O Na Mg P S Cl K Ca Fe I):
  #45 soft tissue (AM_media.dat:49), #49 adipose tissue (:53), #50 lung tissue
  (:54), #2 mineral bone (:6), #53 air (:57, N 80 / O 20 as used for id 0 by
  icrpphantomimportpipeline.cpp:273).  Water is H 0.111894 / O 0.888106 by mass.
Material id 0 is air because the reference zeroes dose where material == 0
(simulationpipeline.cpp:180-185).  Grids are centred on the origin, spacing in
cm, flat index x + nx (y + ny z) (otherphantomimportpipeline.cpp:44).
Geometry draws use numpy's PCG64 generator seeded with the phantom seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ICRP_Z = (1, 6, 7, 8, 11, 12, 15, 16, 17, 19, 20, 26, 53)


def _icrp(percent: tuple) -> dict[int, float]:
    return {z: p / 100.0 for z, p in zip(ICRP_Z, percent) if p > 0}


AIR = {7: 0.80, 8: 0.20}
WATER = {1: 0.111894, 8: 0.888106}
SOFT_TISSUE = _icrp((10.4, 23.1, 2.8, 62.7, 0.1, 0.0, 0.2, 0.3, 0.2, 0.2, 0.0, 0.0, 0.0))
ADIPOSE = _icrp((11.4, 58.8, 0.8, 28.7, 0.1, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.0, 0.0))
LUNG = _icrp((10.3, 10.7, 3.2, 74.6, 0.2, 0.0, 0.2, 0.3, 0.3, 0.2, 0.0, 0.0, 0.0))
BONE = _icrp((3.6, 15.9, 4.2, 44.8, 0.3, 0.2, 9.4, 0.3, 0.0, 0.0, 21.3, 0.0, 0.0))

AIR_DENSITY = 0.001205


@dataclass
class Phantom:
    dims: tuple[int, int, int]
    spacing: tuple[float, float, float]
    density: np.ndarray        # float64, flat
    material: np.ndarray       # uint8, flat
    materials: list[dict[int, float]]
    names: list[str]

    @property
    def size(self) -> int:
        return int(np.prod(self.dims))


def water_cube(n: int = 64, spacing: float = 0.1) -> Phantom:
    """n^3 water (1.0 g/cm3) at `spacing` cm; air is material 0 (unused)."""
    size = n ** 3
    return Phantom((n, n, n), (spacing,) * 3, np.ones(size), np.ones(size, dtype=np.uint8),
                   [AIR, WATER], ["air", "water"])


def _centres(dims, spacing):
    ax = [(np.arange(d) + 0.5 - d / 2.0) * s for d, s in zip(dims, spacing)]
    return ax


def ct_body(dims=(256, 256, 200), spacing=(0.15, 0.15, 0.15), seed: int = 1) -> Phantom:
    """Seeded elliptic body: soft tissue inside an adipose shell, two lungs
    (0.26 g/cm3), spine and ribs of cortical bone (1.92 g/cm3), air outside.

    The body ellipse has full axes 32 x 22 cm times a seeded scale in [0.85, 1]
    clipped to the field of view."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    xs, ys, zs = _centres(dims, spacing)
    fov_x = nx * spacing[0]
    fov_y = ny * spacing[1]
    scale = rng.uniform(0.85, 1.0)
    a = min(16.0 * scale, 0.48 * fov_x)          # semi-axes (cm)
    b = min(11.0 * scale, 0.48 * fov_y)
    shell = rng.uniform(1.0, 2.0)                 # adipose thickness (cm)
    lung_cx = rng.uniform(0.40, 0.50) * a
    lung_a = rng.uniform(0.22, 0.28) * a
    lung_b = rng.uniform(0.50, 0.60) * b
    lung_z0 = zs[0] + rng.uniform(0.25, 0.35) * (zs[-1] - zs[0])
    spine_r = rng.uniform(1.2, 1.8)
    spine_y = -(b - shell - spine_r - rng.uniform(0.5, 1.5))
    rib_pitch = rng.uniform(2.2, 2.8)
    rib_w = 0.6

    X, Y = np.meshgrid(xs, ys, indexing="xy")    # (ny, nx)
    body = (X / a) ** 2 + (Y / b) ** 2 <= 1.0
    inner = (X / (a - shell)) ** 2 + (Y / (b - shell)) ** 2 <= 1.0
    lungs = ((np.abs(X) - lung_cx) / lung_a) ** 2 + (Y / lung_b) ** 2 <= 1.0
    spine = X ** 2 + (Y - spine_y) ** 2 <= spine_r ** 2
    ring_o = (X / (a - shell)) ** 2 + (Y / (b - shell)) ** 2 <= 1.0
    ring_i = (X / (a - shell - 0.8)) ** 2 + (Y / (b - shell - 0.8)) ** 2 <= 1.0
    rib_ring = ring_o & ~ring_i & (Y > spine_y)

    mat = np.zeros((nz, ny, nx), dtype=np.uint8)
    for k, z in enumerate(zs):
        sl = np.zeros((ny, nx), dtype=np.uint8)
        sl[body] = 2                       # adipose
        sl[inner] = 1                      # soft tissue
        if z >= lung_z0:
            sl[lungs & inner] = 3          # lung
            if (z - lung_z0) % rib_pitch < rib_w:
                sl[rib_ring] = 4           # ribs
        sl[spine] = 4                      # spine
        mat[k] = sl
    dens_of = np.array([AIR_DENSITY, 1.03, 0.95, 0.26, 1.92])
    material = mat.ravel()
    density = dens_of[material]
    return Phantom(tuple(dims), tuple(spacing), density, material,
                   [AIR, SOFT_TISSUE, ADIPOSE, LUNG, BONE], ["air", "soft tissue", "adipose", "lung", "bone"])


def ellipsoid(dims, spacing, seed: int = 1) -> Phantom:
    """Seeded ellipsoid of soft tissue with a bone core and a lung-density
    inclusion, air outside (batched uneven workloads)."""
    rng = np.random.default_rng(seed)
    xs, ys, zs = _centres(dims, spacing)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    half = [d * s / 2.0 for d, s in zip(dims, spacing)]
    ax = [h * rng.uniform(0.6, 0.9) for h in half]
    body = (X / ax[0]) ** 2 + (Y / ax[1]) ** 2 + (Z / ax[2]) ** 2 <= 1.0
    core = (X / (0.2 * ax[0])) ** 2 + (Y / (0.2 * ax[1])) ** 2 <= 1.0
    cx = rng.uniform(-0.3, 0.3) * ax[0]
    incl = ((X - cx) / (0.3 * ax[0])) ** 2 + (Y / (0.3 * ax[1])) ** 2 + (Z / (0.4 * ax[2])) ** 2 <= 1.0
    mat = np.zeros(X.shape, dtype=np.uint8)
    mat[body] = 1
    mat[body & incl] = 3
    mat[body & core] = 4
    dens_of = np.array([AIR_DENSITY, 1.03, 0.95, 0.26, 1.92])
    material = mat.ravel()
    return Phantom(tuple(dims), tuple(spacing), dens_of[material], material,
                   [AIR, SOFT_TISSUE, ADIPOSE, LUNG, BONE], ["air", "soft tissue", "adipose", "lung", "bone"])
