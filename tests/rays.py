"""Ray sets for the traversal checks (CPU restatement and GPU bit-exact)."""
from __future__ import annotations

import numpy as np


def random_rays(rng, dims, spacing, n_rays, n_steps):
    """Random rays, a quarter of them axis-aligned and starting on voxel faces
    (ties), half of those with equal steps of 0.25 cm."""
    half = np.array([d * s / 2 for d, s in zip(dims, spacing)], dtype=np.float32)
    pos = rng.uniform(-3, 3, (n_rays, 3)).astype(np.float32) * 2.0
    dirs = rng.normal(size=(n_rays, 3))
    k = n_rays // 4
    dirs[:k] = 0.0
    dirs[np.arange(k), rng.integers(0, 3, k)] = rng.choice([-1.0, 1.0], k)
    pos[:k] = (np.round(rng.uniform(-1, 1, (k, 3)) * 4) / 4 * half).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = dirs.astype(np.float32)
    steps = rng.exponential(0.4, (n_rays, n_steps)).astype(np.float32)
    steps[: k // 2] = np.float32(0.25)
    return pos, dirs, steps


def face_rays(rng, dims, spacing, n_steps, ulps=3):
    """Rays whose first lookup is at a voxel face of one axis or within `ulps`
    FP32 ulps of it (the face b_min + k * spacing in FP32, every k): the start
    point lies in the grid, the first step is 0, so that lookup sees the point
    itself.  Different roundings of the lookup pick different voxels here."""
    f = np.float32
    sp = np.asarray(spacing, dtype=np.float32)
    bmin = [f(-0.5) * (f(dims[a]) * sp[a]) for a in range(3)]
    pos, dirs = [], []
    for a in range(3):
        for k in range(dims[a] + 1):
            x0 = f(bmin[a] + f(k) * sp[a])
            for u in range(-ulps, ulps + 1):
                x = x0
                for _ in range(abs(u)):
                    x = np.nextafter(x, f(np.inf) if u > 0 else f(-np.inf), dtype=np.float32)
                p = (rng.uniform(-0.45, 0.45, 3) * np.array(dims) * sp).astype(np.float32)
                p[a] = x
                d = np.zeros(3, np.float32)
                d[rng.integers(0, 3)] = rng.choice([-1.0, 1.0])
                pos.append(p)
                dirs.append(d)
    pos = np.array(pos, np.float32)
    dirs = np.array(dirs, np.float32)
    steps = rng.exponential(0.4, (len(pos), n_steps)).astype(np.float32)
    steps[:, 0] = 0.0
    return pos, dirs, steps
