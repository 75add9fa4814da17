"""The oracle's DoseScore conversion (simulationpipeline.cpp:174-229; mGy,
datacontainer.hpp:109) pinned by closed forms: 1 keV in 1 g is
1.602176634e-10 mGy; the variance of the sum of H i.i.d. history scores."""
from __future__ import annotations

import numpy as np
import pytest

from physics_refs import MGY_PER_KEV_G, kn_transfer_fraction


def test_dose_unit_closed_form(oracle):
    # 1 keV in a 1 cm3 voxel of density 1 g/cm3: 1.602176634e-16 J / 1e-3 kg = 1.602176634e-13 Gy
    d, v = oracle.dose_scores([1.0, 2.0, 0.0], (1.0, 1.0, 1.0), 1, [1.0, 1.0, 5.0], [1.0, 1.0, 25.0])
    assert d[0] == pytest.approx(1.602176634e-10, rel=1e-15)
    assert d[1] == pytest.approx(0.5 * 1.602176634e-10, rel=1e-15)
    assert d[2] == 0.0 and np.all(v == 0.0)   # massless voxel; one history has no variance
    # a 0.1 cm cube of bone (1.92 g/cm3): mass 1.92e-6 kg; 60 keV -> 60 * 1.602e-16 / 1.92e-6 Gy
    d, _ = oracle.dose_scores([1.92], (0.1, 0.1, 0.1), 10, [60.0], [3600.0])
    assert d[0] == pytest.approx(60 * 1.602176634e-16 / 1.92e-6 * 1e3, rel=1e-13)


def test_dose_variance_is_that_of_a_sum(oracle):
    """For H i.i.d. history scores x_h the estimator is H s^2 (s^2 the sample
    variance), the variance of the sum, converted like the dose squared."""
    rng = np.random.default_rng(4)
    x = rng.exponential(20.0, 1000) * (rng.random(1000) < 0.3)
    H = len(x)
    d, v = oracle.dose_scores([1.0], (1.0, 1.0, 1.0), H, [x.sum()], [np.sum(x * x)])
    k = 1.602176634e-10
    assert d[0] == pytest.approx(x.sum() * k, rel=1e-13)
    assert v[0] == pytest.approx(H * np.var(x, ddof=1) * k * k, rel=1e-10)


def test_kn_transfer_fraction_limits():
    """Thomson limit: a scatter at k -> 0 gives the electron nothing; at 1 MeV
    the mean transfer of a free-electron scatter is 0.440 of the photon energy
    (Attix, Introduction to Radiological Physics and Radiation Dosimetry,
    1986, fig. 7.7 / table of sigma_tr / sigma)."""
    assert kn_transfer_fraction(0.01) < 1e-4
    assert kn_transfer_fraction(1000.0) == pytest.approx(0.440, abs=3e-3)
    assert MGY_PER_KEV_G == pytest.approx(1.602176634e-16 / 1e-3 * 1e3, rel=1e-15)
