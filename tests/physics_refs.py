"""Closed forms used as references by the dose tests."""
from __future__ import annotations

from scipy import integrate

ME_C2 = 510.99895   # keV (CODATA 2018)
MGY_PER_KEV_G = 1.602176634e-10   # 1 keV imparted to 1 g, in mGy (1 eV = 1.602176634e-19 J, exact)


def kn_transfer_fraction(E: float) -> float:
    """Mean fraction of the photon energy given to a free electron in a
    Klein-Nishina scatter: <1 - P> over dsigma/dOmega ~ P^2 (P + 1/P - sin^2),
    P = E'/E = 1 / (1 + k (1 - cos)), k = E / m_e c^2."""
    k = E / ME_C2

    def dcs(c, f):
        P = 1.0 / (1.0 + k * (1.0 - c))
        return P * P * (P + 1.0 / P - (1.0 - c * c)) * f(P)

    num = integrate.quad(lambda c: dcs(c, lambda P: 1.0 - P), -1.0, 1.0, epsabs=0, epsrel=1e-12)[0]
    den = integrate.quad(lambda c: dcs(c, lambda P: 1.0), -1.0, 1.0, epsabs=0, epsrel=1e-12)[0]
    return num / den
