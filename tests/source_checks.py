"""Distribution checks of sampled source particles (ax, ay, direction, E, w)
against what the reference fixes: the collimation geometry of
pointsFromCollimations (beamactorcontainer.cpp:44-69), the discrete
spectrum's weights, and piecewise-linear bowtie interpolation of the
bowtiefilters.json nodes (bowtiefilterreader.cpp:51-97).  Shared by the CPU
tests of the oracle and the GPU tests of the transport kernels' sampling."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
from scipy import stats

BOWTIE_JSON = Path(__file__).resolve().parent.parent / "opendxmc_amd" / "data" / "bowtie_filters.json"


def bowtie_nodes(name: str) -> tuple[np.ndarray, np.ndarray]:
    """(angle, weight) nodes of a filter in the reference's JSON format."""
    doc = json.loads(BOWTIE_JSON.read_text())
    for f in doc["filters"]:
        if f.get("name") == name:
            a = np.array([d["angle"] for d in f["filterdata"]], dtype=np.float64)
            w = np.array([d["weight"] for d in f["filterdata"]], dtype=np.float64)
            return a, w
    raise KeyError(name)


def bowtie_weight(nodes, ang):
    """Linear interpolation in |angle| between the nodes sorted by |angle|,
    clamped at the ends, relative to the weight at angle 0."""
    a, w = nodes
    o = np.argsort(np.abs(a), kind="stable")
    aa, ww = np.abs(a)[o], w[o]
    return np.interp(np.abs(ang), aa, ww) / np.interp(0.0, aa, ww)


def corner_rays(cosines, half):
    """The four field-corner directions of pointsFromCollimations
    (beamactorcontainer.cpp:44-69): dir = cos0 x cos1, signs of sin(x) by
    corner (+, +, -, -), of sin(y) (+, -, -, +)."""
    c0, c1 = np.asarray(cosines[:3], float), np.asarray(cosines[3:], float)
    n = np.cross(c0, c1)
    sx, sy = math.sin(half[0]), math.sin(half[1])
    out = []
    for i, ysign in enumerate((1, -1, -1, 1)):
        x = sx if i < 2 else -sx
        y = sy * ysign
        out.append(c0 * x + c1 * y + n * math.sqrt(1 - x * x - y * y))
    return np.array(out)


def direction_of(cosines, ax, ay):
    """Direction at collimation angles (ax, ay), the same parameterisation as
    the corner rays (one corner per sign pattern of ax, ay)."""
    c0, c1 = np.asarray(cosines[:3], float), np.asarray(cosines[3:], float)
    n = np.cross(c0, c1)
    sx, sy = np.sin(ax)[:, None], np.sin(ay)[:, None]
    d = c0 * sx + c1 * sy + n * np.sqrt(1 - sx * sx - sy * sy)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def check_angles_uniform(ax, ay, half, p_min=1e-3):
    for a, h in ((ax, half[0]), (ay, half[1])):
        if h == 0:
            assert np.all(a == 0)
            continue
        assert np.all(np.abs(a) <= h * (1 + 1e-6))
        p = stats.kstest(a, stats.uniform(loc=-h, scale=2 * h).cdf).pvalue
        assert p > p_min, f"collimation angle not uniform: KS p = {p:.2e}"


def check_energy_chi2(E, energies, weights, p_min=1e-3):
    """Chi-square of the drawn energies against the spectrum weights (bins
    with fewer than 20 expected draws pooled into one)."""
    energies = np.asarray(energies, float)
    prob = np.asarray(weights, float) / np.sum(weights)
    # every draw is one of the spectrum's energies (FP32 on the device):
    # the nearest one, found between the midpoints of sorted energies
    order = np.argsort(energies, kind="stable")
    es = energies[order]
    k = order[np.searchsorted(0.5 * (es[1:] + es[:-1]), E.astype(np.float64))]
    assert np.allclose(energies[k], E, rtol=1e-6, atol=0), "drawn energy not in the spectrum"
    counts = np.bincount(k, minlength=len(energies)).astype(float)
    expected = prob * len(E)
    big = expected >= 20
    obs = np.append(counts[big], counts[~big].sum())
    exp = np.append(expected[big], expected[~big].sum())
    if exp[-1] == 0:
        obs, exp = obs[:-1], exp[:-1]
    chi2 = float(np.sum((obs - exp) ** 2 / exp))
    p = stats.chi2.sf(chi2, len(obs) - 1)
    assert p > p_min, f"energies do not follow the spectrum: chi2 {chi2:.1f} on {len(obs) - 1} dof, p = {p:.2e}"
