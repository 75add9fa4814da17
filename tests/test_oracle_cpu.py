"""The oracle pinned against closed forms, before it is trusted as the checker.

* mixture rule: at physics-list nodes the oracle's mass attenuation equals the
  mass-fraction sum of the element cross sections exactly;
* Woodcock tracking: the uncollided fraction of a pencil beam through a
  layered slab equals Beer-Lambert exp(-sum mu_i rho_i t_i) (a heterogeneous
  grid whose majorant is set by the bone layer, so virtual steps dominate);
* energy bookkeeping: emitted = deposited + escaped for every run;
* samplers: chi-square of sampled cos(theta) against Klein-Nishina, Thomson,
  KN x S(x) and Thomson x F^2(x) densities, and E'/E = 1/(1 + k(1 - cos));
* RNG: splitmix64's published first outputs for seed 0
  (0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4) seed xoshiro128+.
"""
from __future__ import annotations

import math
from fractions import Fraction
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "opendxmc_amd" / "physics"))
import gen_physics_list as g  # noqa: E402

from opendxmc_amd import phantoms  # noqa: E402
from rays import face_rays, random_rays  # noqa: E402

NA = 6.02214076e23
ME_C2 = 510.99895
HC = 12.398419843320026


def test_mixture_rule_at_nodes(oracle):
    e = g.energy_grid()[[40, 120, 200]]
    comp = {1: 0.3, 8: 0.5, 20: 0.2}
    got = oracle.attenuation(comp, e)
    exp_inc = np.zeros(3)
    exp_coh = np.zeros(3)
    for z, w in comp.items():
        inc, coh = g.scatter_cross_sections(z, e)
        exp_inc += w * NA / g.ATOMIC_MASS[z - 1] * 1e-24 * inc
        exp_coh += w * NA / g.ATOMIC_MASS[z - 1] * 1e-24 * coh
    np.testing.assert_allclose(got[:, 1], exp_inc, rtol=1e-9)
    np.testing.assert_allclose(got[:, 2], exp_coh, rtol=1e-9)


def layered_slab():
    # 8 x 8 x 60 voxels of 0.1 cm: water 2 cm, bone 1 cm, lung 3 cm along z
    dims = (8, 8, 60)
    mat = np.zeros(dims[::-1], dtype=np.uint8)
    dens = np.zeros(dims[::-1])
    mat[0:20], dens[0:20] = 1, 1.0
    mat[20:30], dens[20:30] = 2, 1.92
    mat[30:60], dens[30:60] = 3, 0.26
    mats = [phantoms.AIR, phantoms.WATER, phantoms.BONE, phantoms.LUNG]
    return dims, (0.1, 0.1, 0.1), dens.ravel(), mat.ravel(), mats


def pencil(n, energy=60.0, pos=(0.05, 0.05, -10.0)):
    ex = np.zeros(1, dtype=np.dtype([("position", "<f8", 3), ("cosines", "<f8", 6),
                                     ("collimation_half_angles", "<f8", 2), ("weight", "<f8"),
                                     ("spectrum_index", "<u4"), ("bowtie_index", "<i4"), ("n_particles", "<u8")],
                                    align=True))
    ex["position"] = pos
    ex["cosines"] = (1, 0, 0, 0, 1, 0)
    ex["weight"] = 1.0
    ex["bowtie_index"] = -1
    ex["n_particles"] = n
    return ex, [(np.array([energy]), np.array([1.0]))]


@pytest.mark.parametrize("level", [0, 1])
def test_woodcock_beer_lambert(oracle, level):
    dims, sp, dens, mat, mats = layered_slab()
    n = 400_000
    ex, spec = pencil(n)
    _, _, _, c = oracle.transport(dims, sp, dens, mat, mats, ex, spec, level=level, seed=3, threads=8)
    mu = [oracle.attenuation(m, [60.0])[0, 3] for m in mats]
    expected = math.exp(-(mu[1] * 1.0 * 2.0 + mu[2] * 1.92 * 1.0 + mu[3] * 0.26 * 3.0))
    got = c["uncollided"] / n
    se = math.sqrt(expected * (1 - expected) / n)
    assert abs(got - expected) < 4 * se, (got, expected, se)
    assert c["steps"] > 1.5 * c["real_events"]   # majorant from bone: many virtual steps
    assert abs(c["deposited_energy"] + c["escaped_energy"] - c["emitted_energy"]) < 1e-9 * c["emitted_energy"]


def test_energy_bookkeeping_and_tally_sum(oracle):
    ph = phantoms.ct_body((32, 32, 20), (1.2, 1.2, 1.2), seed=2)
    ex, spec = pencil(50_000, energy=80.0, pos=(0.3, 0.2, -30.0))
    se, se2, nev, c = oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, ph.materials, ex, spec,
                                       level=1, seed=4, threads=4)
    assert abs(c["deposited_energy"] + c["escaped_energy"] - c["emitted_energy"]) < 1e-9 * c["emitted_energy"]
    assert se.sum() == pytest.approx(c["deposited_energy"], rel=1e-12)
    assert int(nev.sum()) == c["deposits"]
    assert np.all(se2 >= 0) and np.all(se2 <= se ** 2 + 1e-9)


def chi2_pvalue(samples, pdf, lo=-1.0, hi=1.0, bins=40):
    edges = np.linspace(lo, hi, bins + 1)
    counts, _ = np.histogram(samples, edges)
    fine = np.linspace(lo, hi, bins * 200 + 1)
    dens = pdf(fine)
    cum = np.concatenate([[0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(fine))])
    probs = np.diff(np.interp(edges, fine, cum)) / cum[-1]
    exp = probs * len(samples)
    keep = exp > 5
    chi2 = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    return stats.chi2.sf(chi2, keep.sum() - 1)


def kn_pdf(k):
    def f(mu):
        p = 1 / (1 + k * (1 - mu))
        return p * p * (p + 1 / p - (1 - mu * mu))
    return f


@pytest.mark.parametrize("e", [20.0, 60.0, 150.0])
def test_klein_nishina_sampler(oracle, e):
    c, eps = oracle.sample_angles(phantoms.WATER, e, 0, 200_000, seed=9)
    k = e / ME_C2
    assert chi2_pvalue(c, kn_pdf(k)) > 1e-4
    np.testing.assert_allclose(eps, 1 / (1 + k * (1 - c)), rtol=1e-10)


def test_thomson_sampler(oracle):
    c, _ = oracle.sample_angles(phantoms.WATER, 60.0, 2, 200_000, seed=10)
    assert chi2_pvalue(c, lambda mu: 1 + mu * mu) > 1e-4


@pytest.mark.parametrize("e", [15.0, 40.0])
def test_compton_with_scatter_function(oracle, e):
    comp = phantoms.BONE
    c, _ = oracle.sample_angles(comp, e, 1, 200_000, seed=11)
    k = e / ME_C2

    def pdf(mu):
        x = e / HC * np.sqrt(0.5 * (1 - mu))
        s, _, _ = oracle.scatter_functions(comp, x)
        return kn_pdf(k)(mu) * s
    assert chi2_pvalue(c, pdf) > 1e-4


@pytest.mark.parametrize("e", [20.0, 60.0])
def test_rayleigh_with_form_factor(oracle, e):
    comp = phantoms.SOFT_TISSUE
    c, _ = oracle.sample_angles(comp, e, 3, 200_000, seed=12)

    def pdf(mu):
        x = e / HC * np.sqrt(0.5 * (1 - mu))
        _, f2, _ = oracle.scatter_functions(comp, x)
        return (1 + mu * mu) * f2
    # the density is sharply forward-peaked: bin in cos over [0.5, 1] plus a tail check
    assert chi2_pvalue(c[c > 0.5], pdf, lo=0.5, hi=1.0, bins=50) > 1e-4


def test_rng_known_answer(oracle):
    s = oracle.rng_stream(0, 0, 1)
    # splitmix64(0) -> 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4; xoshiro128+ returns s0 + s3
    assert int(s[0]) == (0x7B1DCDAF + 0x6E789E6A) & 0xFFFFFFFF


def test_traverse_f32_indices_contain_points(oracle):
    rng = np.random.default_rng(1)
    dims = (10, 12, 14)
    sp = np.array([0.3, 0.25, 0.2], dtype=np.float32)
    mat = rng.integers(0, 4, int(np.prod(dims))).astype(np.uint8)
    pos = rng.uniform(-6, 6, (500, 3)).astype(np.float32)
    d = rng.normal(size=(500, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    steps = rng.exponential(0.3, (500, 10)).astype(np.float32)
    idx, mm = oracle.traverse_f32(dims, sp, mat, pos, d, steps)
    half = np.array(dims) * sp.astype(np.float64) / 2
    for r in range(500):
        # entry point in float64, then the steps
        p0 = pos[r].astype(np.float64)
        dd = d[r].astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta, tb = (-half - p0) / dd, (half - p0) / dd
        t0 = max(0.0, np.nanmax(np.minimum(ta, tb)))
        t1 = np.nanmin(np.maximum(ta, tb))
        if t0 > t1:
            assert np.all(idx[r] == -1)
            continue
        p = p0 + t0 * dd + np.cumsum(steps[r].astype(np.float64))[:, None] * dd
        ijk = np.floor((p + half) / sp.astype(np.float64)).astype(int)
        inside = np.all((ijk >= 0) & (ijk < np.array(dims)), axis=1)
        # compare away from faces (float32 rounding decides exact ties)
        frac = (p + half) / sp.astype(np.float64) - np.floor((p + half) / sp.astype(np.float64))
        clear = np.all((frac > 1e-3) & (frac < 1 - 1e-3), axis=1)
        flat = ijk[:, 0] + dims[0] * (ijk[:, 1] + dims[1] * ijk[:, 2])
        exp_idx = np.where(inside, flat, -1)
        np.testing.assert_array_equal(idx[r][clear], exp_idx[clear])
        np.testing.assert_array_equal(mm[r][clear & inside], mat[exp_idx[clear & inside]])


# ----------------------------------------------------------------- level 2
def unit_profile(x):
    return 8.0 / (3.0 * np.pi) * (1.0 + x * x) ** -3


def test_impulse_shells_selection(oracle):
    f, b, p0 = oracle.impulse_shells(phantoms.WATER)
    assert len(f) == 4 and f.sum() == pytest.approx(1.0, abs=1e-12)   # O 1s, O 2s, O 2p, H 1s
    assert b[0] == pytest.approx(g.K_EDGE_KEV[7])                    # oxygen K binding first
    assert f[0] == pytest.approx(2 * 0.888106 / 15.999 / (0.111894 / 1.008 + 8 * 0.888106 / 15.999), rel=1e-9)
    assert p0[0] == pytest.approx(7.7 * 7.2973525693e-3, rel=1e-12)   # Slater 1s charge 8 - 0.3
    f, b, p0 = oracle.impulse_shells(phantoms.BONE)
    assert len(f) == 5 and f.sum() == pytest.approx(1.0, abs=1e-12)
    assert b[0] == pytest.approx(g.K_EDGE_KEV[19]) and np.all(np.diff(b[:4]) <= 0)


def recovered_pz(E, c, eps):
    a = E / ME_C2
    b = a * eps
    return (a * b * (1 - c) - (a - b)) / np.sqrt(a * a + b * b - 2 * a * b * c)


def test_impulse_approximation_profile(oracle):
    """Hydrogen (one shell, 13.6 eV binding, so no kinematic truncation): the
    p_z recovered from (E, E', theta) follows J(p_z) = g(p_z/p0)/p0 with
    p0 = alpha m_e c -- the kinematics and the choice of root."""
    E = 60.0
    c, eps = oracle.sample_angles({1: 1.0}, E, 4, 400_000, seed=13)
    x = recovered_pz(E, c, eps) / 7.2973525693e-3
    # at small angles E' <= E - B truncates the profile (physics, not sampling):
    # compare the large-angle samples, where no truncation occurs
    big = (c < 0.5) & (np.abs(x) < 3.0)
    assert chi2_pvalue(x[big], unit_profile, lo=-3, hi=3, bins=36) > 1e-4
    a = E / ME_C2
    # Doppler broadening leaves the mean near Klein-Nishina: <E'> within 0.1 %
    assert np.mean(a * eps) == pytest.approx(np.mean(a / (1 + a * (1 - c))), rel=1e-3)


def test_impulse_approximation_energy_bound(oracle):
    # water at 20 keV: every scatter leaves at least its shell's binding energy
    E = 20.0
    c, eps = oracle.sample_angles(phantoms.WATER, E, 4, 100_000, seed=16)
    f, b, p0 = oracle.impulse_shells(phantoms.WATER)
    assert np.all(eps > 0) and np.all(E - eps * E >= b.min() - 1e-9)
    # the recovered |p_z| never exceeds what the inverse-CDF range allows
    assert np.all(np.abs(recovered_pz(E, c, eps)) <= 8.0 * p0.max() * (1 + 1e-9))


def test_fluorescence_yield_and_lines(oracle):
    """Pure iodine at 50 keV: P(line) = sigma_K / sigma_pe x omega_K, lines Ka1, Ka2, Kb."""
    E = 50.0
    _, eps = oracle.sample_angles({53: 1.0}, E, 5, 400_000, seed=14)
    lines = eps * E
    pe_e, pe = g.photo_table(53, g.energy_grid())
    pk = g.photo_k_table(53, pe_e)
    frac_k = np.exp(np.interp(np.log(E), np.log(pe_e), np.log(pk))) / np.exp(np.interp(np.log(E), np.log(pe_e), np.log(pe)))
    omega, e1, e2, e3, p1, p2, p3 = g.k_lines(53)
    expected = frac_k * omega
    got = np.mean(lines > 0)
    assert got == pytest.approx(expected, abs=4 * math.sqrt(expected * (1 - expected) / len(lines)))
    emitted = lines[lines > 0]
    assert set(np.round(np.unique(emitted), 6)) <= {round(e1, 6), round(e2, 6), round(e3, 6)}
    share = np.array([np.mean(np.isclose(emitted, e)) for e in (e1, e2, e3)])
    np.testing.assert_allclose(share, [p1, p2, p3], atol=0.01)
    # below the K edge nothing fluoresces above 1 keV except L lines (not tracked)
    _, eps = oracle.sample_angles({53: 1.0}, 30.0, 5, 20_000, seed=15)
    assert np.all(eps == 0)


def test_level2_energy_bookkeeping(oracle):
    ph = phantoms.ct_body((32, 32, 20), (1.2, 1.2, 1.2), seed=2)
    mats = list(ph.materials)
    mats[4] = {53: 0.05, **{z: w * 0.95 for z, w in phantoms.SOFT_TISSUE.items()}}   # iodinated insert
    ex, spec = pencil(50_000, energy=80.0, pos=(0.3, 0.2, -30.0))
    se, _, nev, c = oracle.transport(ph.dims, ph.spacing, ph.density, ph.material, mats, ex, spec,
                                     level=2, seed=4, threads=4)
    assert abs(c["deposited_energy"] + c["escaped_energy"] - c["emitted_energy"]) < 1e-9 * c["emitted_energy"]
    assert se.sum() == pytest.approx(c["deposited_energy"], rel=1e-12)


# ------------------------------------------------ traversal, exact restatement
def _round_f32(x: Fraction) -> np.float32:
    """Round an exact rational to the nearest FP32 value (ties to even)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    # scale so that 2^23 <= m < 2^24
    while True:
        m = a / Fraction(2) ** (e - 23) if e >= 23 else a * Fraction(2) ** (23 - e)
        if m < 2 ** 23:
            e -= 1
        elif m >= 2 ** 24:
            e += 1
        else:
            break
    q, r = divmod(m.numerator, m.denominator)
    if 2 * r > m.denominator or (2 * r == m.denominator and q % 2 == 1):
        q += 1
    return np.float32(sign * q * 2.0 ** (e - 23))


def _fma_f32(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def _traverse_exact(dims, sp, mat, pos, dirs, steps):
    """Independent restatement of the tracking's voxel lookup: every FP32
    operation is evaluated exactly in rationals and rounded once (numpy float32
    for the plain operations, _round_f32 for the fused multiply-adds)."""
    f = np.float32
    bmin = [f(-0.5) * (f(dims[a]) * sp[a]) for a in range(3)]
    bmax = [-b for b in bmin]
    inv = [f(1.0) / sp[a] for a in range(3)]
    off = [-bmin[a] * inv[a] for a in range(3)]
    n_rays, n_steps = steps.shape
    idx = np.full((n_rays, n_steps), -1, dtype=np.int64)
    mm = np.full((n_rays, n_steps), -1, dtype=np.int32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        for r in range(n_rays):
            p = [pos[r, a] for a in range(3)]
            d = [dirs[r, a] for a in range(3)]
            t0, t1 = f(0.0), f(3.0e38)
            for a in range(3):
                i_d = f(1.0) / d[a]
                ta, tb = (bmin[a] - p[a]) * i_d, (bmax[a] - p[a]) * i_d
                t0 = np.fmax(t0, np.fmin(ta, tb))
                t1 = np.fmin(t1, np.fmax(ta, tb))
            if not t0 <= t1:
                continue
            p = [_fma_f32(t0, d[a], p[a]) for a in range(3)]
            for k in range(n_steps):
                p = [_fma_f32(steps[r, k], d[a], p[a]) for a in range(3)]
                c = [int(np.floor(_fma_f32(p[a], inv[a], off[a]))) for a in range(3)]
                if all(0 <= c[a] < dims[a] for a in range(3)):
                    idx[r, k] = c[0] + dims[0] * (c[1] + dims[1] * c[2])
                    mm[r, k] = mat[idx[r, k]]
    return idx, mm


def test_traverse_f32_exact_restatement(oracle):
    rng = np.random.default_rng(11)
    dims = (13, 9, 7)
    sp = np.array([0.25, 0.5, 0.3], dtype=np.float32)
    mat = rng.integers(0, 3, int(np.prod(dims))).astype(np.uint8)
    pos, dirs, steps = random_rays(rng, dims, sp, 160, 12)
    idx, mm = oracle.traverse_f32(dims, sp, mat, pos, dirs, steps)
    eidx, emm = _traverse_exact(dims, sp, mat, pos, dirs, steps)
    assert (idx >= 0).sum() > 160
    np.testing.assert_array_equal(idx, eidx)
    np.testing.assert_array_equal(mm, emm)


@pytest.mark.parametrize("dims,spacing", [((13, 9, 7), (0.25, 0.5, 0.3)), ((64, 64, 64), (0.1, 0.1, 0.1))])
def test_traverse_f32_face_points(oracle, dims, spacing):
    """Lookups at and within 3 ulps of every voxel face: the oracle's choice of
    voxel is the single-rounding fma restated exactly, and differs from both
    an unfused (x * inv + off) and a (x - box_min) * inv lookup on some of
    these points, so either of those mistakes fails here."""
    rng = np.random.default_rng(5)
    sp = np.array(spacing, dtype=np.float32)
    mat = rng.integers(0, 3, int(np.prod(dims))).astype(np.uint8)
    pos, dirs, steps = face_rays(rng, dims, sp, 2)
    idx, mm = oracle.traverse_f32(dims, sp, mat, pos, dirs, steps)
    eidx, emm = _traverse_exact(dims, sp, mat, pos, dirs, steps)
    np.testing.assert_array_equal(idx, eidx)
    np.testing.assert_array_equal(mm, emm)
    # the alternative roundings pick other voxels on some of these points
    f = np.float32
    bmin = [f(-0.5) * (f(dims[a]) * sp[a]) for a in range(3)]
    inv = [f(1.0) / sp[a] for a in range(3)]
    off = [-bmin[a] * inv[a] for a in range(3)]
    n_unfused = n_shifted = 0
    for i in range(len(pos)):
        for a in range(3):
            x = pos[i, a]
            fused = int(np.floor(_fma_f32(x, inv[a], off[a])))
            n_unfused += fused != int(np.floor(f(x * inv[a]) + off[a]))
            n_shifted += fused != int(np.floor((x - bmin[a]) * inv[a]))
    assert n_unfused > 0 and n_shifted > 0


def test_round_f32_matches_numpy():
    rng = np.random.default_rng(2)
    for x in rng.normal(size=200) * 10.0 ** rng.integers(-8, 8, 200):
        assert _round_f32(Fraction(float(x))) == np.float32(x)
    # halfway cases round to even
    one = Fraction(1)
    assert _round_f32(one + Fraction(1, 2 ** 24)) == np.float32(1.0)
    assert _round_f32(one + Fraction(3, 2 ** 24)) == np.float32(1.0) + np.float32(2.0 ** -22)
