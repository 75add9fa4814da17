"""Generate the photon physics list (element cross sections, F(x,Z), S(x,Z), shells).

DXMClib ships an EPICS/Livermore-derived "physics list" next to the executable
(src/app/CMakeLists.txt:38, ``dxmclib_add_physics_list``) and builds every
``Material<5>`` from it (simulationpipeline.cpp:133-145).  That file is not
available offline, so this script writes an equivalent list from published
analytic models.  Every number below comes from a named model; nothing is
fitted to tabulated data.

Models
------
* Electron configuration: Madelung (aufbau) filling order, neutral ground state.
* Orbital screening: Slater's rules (J. C. Slater, Phys. Rev. 36, 57 (1930)),
  giving each subshell an effective charge Z_eff and effective principal
  quantum number n*.  Orbitals are Slater-type, R(r) = r^(n*-1) exp(-Z_eff r / n*).
* Atomic form factor F(x,Z): sum over subshells of the closed-form Fourier
  transform of the Slater-orbital density (exact for hydrogen).
* Incoherent scattering function S(x,Z): Waller-Hartree form without the
  off-diagonal exchange terms, S = Z - sum_i N_i f_i(x)^2.
* Incoherent cross section: Klein-Nishina x S(x,Z) integrated over angle.
* Coherent cross section: Thomson x F(x,Z)^2 integrated over angle.
* Photoelectric cross section: screened hydrogenic model.  Each subshell
  (n, l, N electrons, Slater charge Zs, binding B) contributes
  N/2 * n^-3 * sigma_K(E - B; Zs) * (I/(I+E-B))^l, where sigma_K is Stobbe's
  exact non-relativistic hydrogenic 1s (pair) cross section evaluated at the
  photo-electron kinetic energy E - B and I = Zs^2 Ry, times (Z/Zs)^2 because
  the outgoing electron sees the bare nucleus (Hall's short-distance argument).
  For hydrogen this is Stobbe's formula itself (6.30e-18 cm^2 at threshold).
* Binding energies: measured K edges for Z = 1..92 and L1/L2/L3 edges at
  anchor elements (X-ray edge tables of Bearden & Burr, Rev. Mod. Phys. 39,
  125 (1967)); L edges between anchors are interpolated in sqrt(B); M and
  outer shells are scaled from L3 with Slater charges.
* Atomic masses: IUPAC standard atomic weights.

Output format (little endian) -- read by the product's world builder and by the
oracle:  magic "DXGPHYS2"; u32 zmax, n_energy, n_x; f64 energy[n_energy] (keV);
f64 x[n_x] (1/Angstrom); then for Z = 1..zmax: f64 A (g/mol); f64
incoh[n_energy], coh[n_energy] (barn/atom); f64 F[n_x], S[n_x]; u32 n_pe;
f64 pe_energy[n_pe], pe[n_pe] (barn/atom, energies ascending, an edge appears
twice: value below then value above), pe_k[n_pe] (the K-shell part); u32
n_shell; per shell f64 n, l, occupancy, binding (keV), zeff, n*; then f64
omega_K, E_Ka1, E_Ka2, E_Kb (keV), p_Ka1, p_Ka2, p_Kb.

Correction level 2 (impulse approximation) uses each shell's Compton profile
in the hydrogenic 1s form J(p_z) = 8/(3 pi p0) (1 + (p_z/p0)^2)^-3 with the
shell's momentum scale p0 = zeff/n* (atomic units): the exact profile for a
1s orbital and, for other shells, the profile with the shell's virial mean
kinetic energy.
"""
from __future__ import annotations

import math
import struct
import sys
from pathlib import Path

import numpy as np

# CODATA 2018 values (scipy.constants carries the same numbers).
R_E_CM = 2.8179403262e-13          # classical electron radius [cm]
A0_CM = 5.29177210903e-9           # Bohr radius [cm]
A0_ANG = 0.529177210903            # Bohr radius [Angstrom]
ALPHA = 7.2973525693e-3            # fine-structure constant
RY_KEV = 13.605693122994e-3        # Rydberg energy [keV]
HC_KEV_ANG = 12.398419843320026    # h*c [keV*Angstrom]
ME_C2_KEV = 510.99895000           # electron rest energy [keV]
BARN = 1e-24                       # cm^2

ZMAX = 92
E_MIN_KEV = 1.0
E_MAX_KEV = 200.0
N_ENERGY = 240
N_X = 220

# IUPAC standard atomic weights, Z = 1..92 (g/mol).
ATOMIC_MASS = [
    1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999, 18.998403163, 20.1797,
    22.98976928, 24.305, 26.9815385, 28.085, 30.973761998, 32.06, 35.45, 39.948, 39.0983, 40.078,
    44.955908, 47.867, 50.9415, 51.9961, 54.938044, 55.845, 58.933194, 58.6934, 63.546, 65.38,
    69.723, 72.630, 74.921595, 78.971, 79.904, 83.798, 85.4678, 87.62, 88.90584, 91.224,
    92.90637, 95.95, 98.0, 101.07, 102.90550, 106.42, 107.8682, 112.414, 114.818, 118.710,
    121.760, 127.60, 126.90447, 131.293, 132.90545196, 137.327, 138.90547, 140.116, 140.90766, 144.242,
    145.0, 150.36, 151.964, 157.25, 158.92535, 162.500, 164.93033, 167.259, 168.93422, 173.045,
    174.9668, 178.49, 180.94788, 183.84, 186.207, 190.23, 192.217, 195.084, 196.966569, 200.592,
    204.38, 207.2, 208.98040, 209.0, 210.0, 222.0, 223.0, 226.0, 227.0, 232.0377,
    231.03588, 238.02891,
]

# Measured K-shell binding energies (keV), Z = 1..92.
K_EDGE_KEV = [
    0.0136, 0.0246, 0.0548, 0.1115, 0.188, 0.2838, 0.4016, 0.5320, 0.6854, 0.8701,
    1.0721, 1.3050, 1.5596, 1.8389, 2.1455, 2.4720, 2.8224, 3.2029, 3.6074, 4.0381,
    4.4928, 4.9664, 5.4651, 5.9892, 6.5390, 7.1120, 7.7089, 8.3328, 8.9789, 9.6586,
    10.367, 11.103, 11.867, 12.658, 13.474, 14.326, 15.200, 16.105, 17.038, 17.998,
    18.986, 20.000, 21.044, 22.117, 23.220, 24.350, 25.514, 26.711, 27.940, 29.200,
    30.491, 31.814, 33.169, 34.561, 35.985, 37.441, 38.925, 40.443, 41.991, 43.569,
    45.184, 46.834, 48.519, 50.239, 51.996, 53.789, 55.618, 57.486, 59.390, 61.332,
    63.314, 65.351, 67.416, 69.525, 71.676, 73.871, 76.111, 78.395, 80.725, 83.102,
    85.530, 88.005, 90.526, 93.105, 95.730, 98.404, 101.137, 103.922, 106.755, 109.651,
    112.601, 115.606,
]

# Measured L1, L2, L3 binding energies (keV) at anchor elements.
L_EDGES_KEV = {
    11: (0.0633, 0.0311, 0.0311), 12: (0.0889, 0.0497, 0.0496), 13: (0.1177, 0.0729, 0.0727),
    14: (0.1487, 0.0995, 0.0989), 15: (0.1893, 0.1362, 0.1353), 16: (0.2292, 0.1654, 0.1642),
    17: (0.2702, 0.2016, 0.2000), 18: (0.3263, 0.2506, 0.2484), 19: (0.3771, 0.2965, 0.2940),
    20: (0.4384, 0.3500, 0.3463), 22: (0.5604, 0.4601, 0.4538), 26: (0.8461, 0.7199, 0.7068),
    29: (1.0967, 0.9523, 0.9327), 30: (1.1936, 1.0428, 1.0197), 35: (1.782, 1.596, 1.550),
    38: (2.216, 2.007, 1.940), 42: (2.866, 2.625, 2.520), 47: (3.806, 3.524, 3.351),
    48: (4.018, 3.727, 3.538), 50: (4.465, 4.156, 3.929), 53: (5.188, 4.852, 4.557),
    54: (5.453, 5.107, 4.786), 56: (5.989, 5.624, 5.247), 64: (8.376, 7.930, 7.243),
    74: (12.100, 11.544, 10.207), 78: (13.880, 13.273, 11.564), 79: (14.353, 13.734, 11.919),
    82: (15.861, 15.200, 13.035), 83: (16.388, 15.711, 13.419), 92: (21.757, 20.948, 17.166),
}

MADELUNG = [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (3, 2), (4, 1), (5, 0), (4, 2),
            (5, 1), (6, 0), (4, 3), (5, 2), (6, 1), (7, 0), (5, 3), (6, 2), (7, 1)]
N_STAR = {1: 1.0, 2: 2.0, 3: 3.0, 4: 3.7, 5: 4.0, 6: 4.2, 7: 4.3}


def configuration(z: int) -> list[tuple[int, int, int]]:
    """Ground-state subshell occupancies [(n, l, N)] by Madelung filling."""
    left = z
    conf = []
    for n, l in MADELUNG:
        if left <= 0:
            break
        cap = 2 * (2 * l + 1)
        occ = min(cap, left)
        conf.append((n, l, occ))
        left -= occ
    return conf


def _slater_group(n: int, l: int) -> tuple[int, int]:
    return (n, 0 if l <= 1 else l - 1)   # (n, 0)=sp, (n, 1)=d, (n, 2)=f


def slater_zeff(z: int, conf, n: int, l: int) -> float:
    """Slater's-rules effective charge seen by an electron in subshell (n, l)."""
    g = _slater_group(n, l)
    shield = 0.0
    for (n2, l2, occ) in conf:
        g2 = _slater_group(n2, l2)
        cnt = occ - (1 if (n2, l2) == (n, l) else 0)
        if cnt <= 0:
            continue
        if g2 == g:
            shield += (0.30 if n == 1 else 0.35) * cnt
        elif g2 < g:
            if g[1] == 0:   # s,p electron: 0.85 for shell n-1, 1.00 deeper
                if n2 == n - 1:
                    shield += 0.85 * cnt
                elif n2 <= n - 2:
                    shield += 1.00 * cnt
                # same-n d/f groups sort after sp, so they never reach here
            else:           # d or f electron: everything to the left screens fully
                shield += 1.00 * cnt
    return z - shield


def _interp_l_edges(z: int) -> tuple[float, float, float]:
    if z in L_EDGES_KEV:
        return L_EDGES_KEV[z]
    zs = sorted(L_EDGES_KEV)
    if z < zs[0]:
        return None
    if z > zs[-1]:
        lo, hi = zs[-2], zs[-1]
    else:
        hi = next(k for k in zs if k > z)
        lo = max(k for k in zs if k < z)
    t = (z - lo) / (hi - lo)
    out = []
    for i in range(3):
        a, b = math.sqrt(L_EDGES_KEV[lo][i]), math.sqrt(L_EDGES_KEV[hi][i])
        out.append((a + t * (b - a)) ** 2)
    return tuple(out)


def shells(z: int):
    """Subshell list [(n, l, N, binding_keV, zeff)] for element z."""
    conf = configuration(z)
    ledges = _interp_l_edges(z)
    zeff_l3 = None
    out = []
    for (n, l, occ) in conf:
        zeff = max(slater_zeff(z, conf, n, l), 1.0)
        ns = N_STAR[n]
        if n == 1:
            b = K_EDGE_KEV[z - 1]
        elif n == 2 and ledges is not None:
            b = ledges[0] if l == 0 else ledges[2]
        else:
            b = RY_KEV * (zeff / ns) ** 2
            if ledges is not None and n >= 3:
                # scale the measured L3 edge with the Slater charge ratio
                zl = max(slater_zeff(z, conf, 2, 1), 1.0)
                b = ledges[2] * ((zeff / ns) ** 2) / ((zl / 2.0) ** 2)
        out.append((n, l, occ, max(b, 1e-5), zeff))
        if n == 2 and l == 1:
            zeff_l3 = zeff
    del zeff_l3
    return out


def orbital_ff(q_au: np.ndarray, nstar: float, zeff: float) -> np.ndarray:
    """Form factor of one Slater-orbital electron; q in inverse Bohr radii."""
    a = 2.0 * zeff / nstar
    m = 2.0 * nstar
    q = np.asarray(q_au, dtype=np.float64)
    out = np.ones_like(q)
    nz = q > 1e-12
    qq = q[nz]
    phi = np.arctan2(qq, a)
    # a^(m+1) sin(m phi) / (m q (a^2+q^2)^(m/2)), written in terms of ratios
    out[nz] = np.sin(m * phi) / (m * (qq / a)) * (1.0 + (qq / a) ** 2) ** (-m / 2.0)
    return out


def form_factors(z: int, x: np.ndarray):
    """F(x,Z) and S(x,Z); x = sin(theta/2)/lambda in 1/Angstrom."""
    q = 4.0 * math.pi * x * A0_ANG
    F = np.zeros_like(x)
    sumsq = np.zeros_like(x)
    conf = configuration(z)
    for (n, l, occ) in conf:
        zeff = max(slater_zeff(z, conf, n, l), 1.0)
        f = orbital_ff(q, N_STAR[n], zeff)
        F += occ * f
        sumsq += occ * f * f
    S = z - sumsq
    S = np.clip(S, 0.0, float(z))
    return F, S


def stobbe_pair(eps_kev: np.ndarray, zs: float) -> np.ndarray:
    """Stobbe's hydrogenic 1s photo-ionisation cross section for a pair of
    electrons (cm^2), at photo-electron kinetic energy eps, nuclear charge zs."""
    eps = np.asarray(eps_kev, dtype=np.float64)
    out = np.zeros_like(eps)
    ok = eps > 0
    ih = zs * zs * RY_KEV
    e = eps[ok]
    eta = np.sqrt(ih / e)
    ratio = ih / (ih + e)
    pref = 2.0 * (2.0 ** 9) * math.pi ** 2 * ALPHA * A0_CM ** 2 / (3.0 * zs * zs)
    arccot = np.arctan2(1.0, eta)
    out[ok] = pref * ratio ** 4 * np.exp(-4.0 * eta * arccot) / (-np.expm1(-2.0 * math.pi * eta))
    return out


def photo_subshell(e_kev: np.ndarray, n: int, l: int, occ: int, b: float, zs: float,
                   z: float) -> np.ndarray:
    """Screened hydrogenic subshell cross section (barn/atom).

    The bound density follows the screened charge zs while the outgoing
    electron feels the bare nucleus at the short distances that carry the
    momentum transfer, so the hydrogenic result is scaled by (z/zs)^2."""
    eps = np.asarray(e_kev, dtype=np.float64) - b
    sig = stobbe_pair(eps, zs) * (occ / 2.0) / (n ** 3) * (z / zs) ** 2
    ih = zs * zs * RY_KEV
    if l > 0:
        sig = sig * (ih / (ih + np.maximum(eps, 0.0))) ** l
    return sig / BARN


def photo_element(z: int, e_kev: np.ndarray) -> np.ndarray:
    tot = np.zeros_like(np.asarray(e_kev, dtype=np.float64))
    for (n, l, occ, b, zs) in shells(z):
        tot += photo_subshell(e_kev, n, l, occ, b, zs, z)
    return tot


def kn_dcs(k: float, mu: np.ndarray) -> np.ndarray:
    """Klein-Nishina dsigma/dOmega per electron (cm^2/sr); k = E/mc^2."""
    p = 1.0 / (1.0 + k * (1.0 - mu))
    return 0.5 * R_E_CM ** 2 * p * p * (p + 1.0 / p - (1.0 - mu * mu))


def kn_total(k: float) -> float:
    """Closed-form Klein-Nishina total cross section per electron (cm^2)."""
    a = 1.0 + 2.0 * k
    l = math.log(a)
    return 2.0 * math.pi * R_E_CM ** 2 * (
        (1.0 + k) / k ** 2 * (2.0 * (1.0 + k) / a - l / k) + l / (2.0 * k) - (1.0 + 3.0 * k) / a ** 2)


def _angular_nodes(n_panels: int = 96, order: int = 8):
    """Quadrature on v = sin(theta/2) in [0,1], panels geometric near 0."""
    xg, wg = np.polynomial.legendre.leggauss(order)
    edges = np.concatenate([[0.0], np.geomspace(1e-5, 1.0, n_panels)])
    v, w = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        v.append(0.5 * (b - a) * xg + 0.5 * (b + a))
        w.append(0.5 * (b - a) * wg)
    return np.concatenate(v), np.concatenate(w)


def scatter_cross_sections(z: int, e_kev: np.ndarray):
    """(incoherent, coherent) in barn/atom on the energy grid."""
    v, wv = _angular_nodes()
    inc = np.zeros_like(e_kev)
    coh = np.zeros_like(e_kev)
    mu = 1.0 - 2.0 * v * v               # cos(theta)
    jac = 2.0 * math.pi * 4.0 * v * wv   # dOmega = 2 pi dmu = 2 pi 4 v dv
    for i, e in enumerate(e_kev):
        xm = e / HC_KEV_ANG
        F, S = form_factors(z, xm * v)
        k = e / ME_C2_KEV
        inc[i] = np.sum(kn_dcs(k, mu) * S * jac)
        coh[i] = np.sum(0.5 * R_E_CM ** 2 * (1.0 + mu * mu) * F * F * jac)
    return inc / BARN, coh / BARN


def energy_grid():
    return np.geomspace(E_MIN_KEV, E_MAX_KEV, N_ENERGY)


def x_grid():
    return np.concatenate([[0.0], np.geomspace(1e-3, 40.0, N_X - 1)])


def photo_table(z: int, egrid: np.ndarray):
    """Photoelectric table with each edge inside the grid range written twice."""
    edges = sorted({b for (_, _, _, b, _) in shells(z) if E_MIN_KEV < b < E_MAX_KEV})
    pts = [(e, None) for e in egrid if all(abs(e - b) > 1e-9 for b in edges)]
    es, vals = [], []
    merged = sorted([e for e, _ in pts] + edges)
    for e in merged:
        if any(abs(e - b) < 1e-12 for b in edges):
            below = photo_element(z, np.array([e * (1.0 - 1e-9)]))[0]
            above = photo_element(z, np.array([e * (1.0 + 1e-9)]))[0]
            es += [e, e]
            vals += [below, above]
        else:
            es.append(e)
            vals.append(photo_element(z, np.array([e]))[0])
    return np.array(es), np.array(vals)


def photo_k_table(z: int, pe_e: np.ndarray) -> np.ndarray:
    """K-shell part of the photoelectric table at the nodes of photo_table:
    zero below the K edge (the first of the doubled K-edge nodes)."""
    n, l, occ, b, zs = shells(z)[0]
    out = photo_subshell(pe_e * (1.0 + 1e-9), n, l, occ, b, zs, z)
    dup = np.where((np.abs(pe_e - b) < 1e-12))[0]
    if len(dup) == 2:
        out[dup[0]] = 0.0
    return out


def fluorescence_yield_k(z: int) -> float:
    """K-shell fluorescence yield from the semi-empirical fit
    (omega/(1-omega))^(1/4) = 0.015 + 0.0327 Z - 0.64e-6 Z^3 (Burhop; Bambynek et al. 1972)."""
    q = (0.015 + 0.0327 * z - 0.64e-6 * z ** 3) ** 4
    return q / (1.0 + q)


def k_lines(z: int):
    """(omega_K, E_Ka1, E_Ka2, E_Kb, p_Ka1, p_Ka2, p_Kb): line energies as
    differences of binding energies (K-L3, K-L2, K-M3), Ka2/Ka1 = 1/2 (2p1/2 to
    2p3/2 occupancy) and Kb/Ka rising from 0.12 at Z = 20 to 0.23 at Z = 80."""
    sh = shells(z)
    bk = sh[0][3]
    led = _interp_l_edges(z)
    if led is None or z < 11:
        return (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    ka1 = bk - led[2]
    ka2 = bk - led[1]
    m3 = [b for (n, l, occ, b, zs) in sh if n == 3 and l == 1]
    kb = bk - m3[0] if m3 else 0.0
    rb = float(np.interp(z, [11, 20, 50, 80], [0.0, 0.12, 0.20, 0.23])) if m3 else 0.0
    w1, w2, wb = 1.0, 0.5, 1.5 * rb
    s = w1 + w2 + wb
    return (fluorescence_yield_k(z), ka1, ka2, kb, w1 / s, w2 / s, wb / s)


def write(path: Path, zmax: int = ZMAX) -> None:
    eg = energy_grid()
    xg = x_grid()
    with open(path, "wb") as f:
        f.write(b"DXGPHYS2")
        f.write(struct.pack("<III", zmax, len(eg), len(xg)))
        f.write(eg.astype("<f8").tobytes())
        f.write(xg.astype("<f8").tobytes())
        for z in range(1, zmax + 1):
            inc, coh = scatter_cross_sections(z, eg)
            F, S = form_factors(z, xg)
            pe_e, pe = photo_table(z, eg)
            sh = shells(z)
            f.write(struct.pack("<d", ATOMIC_MASS[z - 1]))
            f.write(inc.astype("<f8").tobytes())
            f.write(coh.astype("<f8").tobytes())
            f.write(F.astype("<f8").tobytes())
            f.write(S.astype("<f8").tobytes())
            f.write(struct.pack("<I", len(pe_e)))
            f.write(pe_e.astype("<f8").tobytes())
            f.write(pe.astype("<f8").tobytes())
            f.write(photo_k_table(z, pe_e).astype("<f8").tobytes())
            f.write(struct.pack("<I", len(sh)))
            for (n, l, occ, b, zs) in sh:
                f.write(struct.pack("<6d", n, l, occ, b, zs, N_STAR[n]))
            f.write(struct.pack("<7d", *k_lines(z)))


def default_path() -> Path:
    return Path(__file__).resolve().parent / "physics_list.bin"


if __name__ == "__main__":
    out = Path(sys.argv[1]) if len(sys.argv) > 1 else default_path()
    write(out)
    print(f"wrote {out} ({out.stat().st_size} bytes)")
