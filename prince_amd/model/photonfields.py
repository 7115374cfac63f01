"""Target photon fields n(E, z) in 1/(GeV cm^3).

* ``CMBPhotonSpectrum``: Planck spectrum at T = 2.725 (1+z) K, evaluated only
  where E/(E_CMB (1+z)) < 709 (reference ``photonfields.py:93-127``).
* ``FlatPhotonSpectrum``: 1e12 (1+z)^2, for tests (``photonfields.py:74-90``).
* ``CombinedPhotonField``: sum of components (``photonfields.py:27-71``).
* ``SyntheticEBL``: extragalactic background light built from two modified
  blackbodies (stellar peak near 1 um, dust peak near 100 um) with seeded
  amplitudes and a fixed redshift scaling (1+z)^3 (1+z)^-1.2.  The reference
  reads measured EBL tables; this field replaces them with a closed form that
  has the same two-peak structure and is linear in its amplitude ``scale``.
"""

from __future__ import annotations

import numpy as np

from . import config

#: 1/(pi^2 (hbar c)^3) in GeV^-3 cm^-3 (photonfields.py:114).
PLANCK_PREF = 1.31868e40


class PhotonField:
    def get_photon_density(self, E, z):  # pragma: no cover - interface
        raise NotImplementedError


class CMBPhotonSpectrum(PhotonField):
    def get_photon_density(self, E, z):
        E = np.asarray(E, dtype=np.float64)
        pref = PLANCK_PREF
        Ered = E / (1.0 + z)
        eratio = Ered / config.E_CMB
        exp_range = eratio < 709.0
        nlocal = np.zeros_like(Ered)
        pref *= (1.0 + z) ** 2
        nlocal[exp_range] = pref * Ered[exp_range] ** 2 / (np.exp(eratio[exp_range]) - 1.0)
        return nlocal


class FlatPhotonSpectrum(PhotonField):
    def get_photon_density(self, E, z):
        nlocal = np.ones_like(np.asarray(E, dtype=np.float64)) * 1e12
        return (1.0 + z) ** 2 * nlocal


class CombinedPhotonField(PhotonField):
    """Sum of photon-field components (instances or classes)."""

    def __init__(self, components):
        self.model_list = []
        for c in components:
            if isinstance(c, (tuple, list)):
                cls, args = c[0], c[1:]
                self.model_list.append(cls(*args))
            elif isinstance(c, type):
                self.model_list.append(c())
            else:
                self.model_list.append(c)

    def add_model(self, model):
        self.model_list.append(model)

    def get_photon_density(self, E, z):
        E = np.atleast_1d(np.asarray(E, dtype=np.float64))
        return np.sum([m.get_photon_density(E, z) for m in self.model_list], axis=0)


class SyntheticEBL(PhotonField):
    """Two diluted modified blackbodies, n(E,z) = scale (1+z)^1.8 n0(E).

    n0(E) = sum_i A_i w_i PLANCK_PREF E^2 (E/kT_i)^beta_i / (exp(E/kT_i) - 1)

    with kT_star = 0.32 eV (E^2 n peak near 1 um), kT_dust = 3.2 meV (peak
    near 100 um), dilution factors w_i that give present-day energy densities
    of roughly 6e-3 and 3e-3 eV/cm^3, beta = (0, 1.5), and amplitudes A_i
    drawn U[0.5, 1.5] from ``numpy.random.default_rng(seed)``.  The density
    is zero where E/kT_i >= 709 (same overflow guard as the CMB).
    """

    KT = (0.32e-9, 3.2e-12)  # GeV
    BETA = (0.0, 1.5)
    DILUTION = (7.0e-15, 2.0e-7)
    Z_POWER = 3.0 - 1.2

    def __init__(self, seed=0, amplitudes=None, scale=1.0):
        if amplitudes is None:
            rng = np.random.default_rng(seed)
            amplitudes = rng.uniform(0.5, 1.5, size=2)
        self.amplitudes = np.asarray(amplitudes, dtype=np.float64)
        self.scale = float(scale)

    def density_z0(self, E):
        E = np.asarray(E, dtype=np.float64)
        out = np.zeros_like(E)
        for amp, kT, beta, w in zip(self.amplitudes, self.KT, self.BETA, self.DILUTION):
            r = E / kT
            ok = r < 709.0
            out[ok] += (
                amp * w * PLANCK_PREF * E[ok] ** 2 * r[ok] ** beta / (np.exp(r[ok]) - 1.0)
            )
        return out

    def get_photon_density(self, E, z):
        return self.scale * (1.0 + z) ** self.Z_POWER * self.density_z0(E)
