"""Cosmic-ray source models: injection spectra and source evolution.

Restates ``cr_sources.py:14-226`` of the reference.  A source holds a
per-species ``injection_grid`` on the state vector and returns

    injection_rate(z) = norm * evolution(z) * injection_grid

(``cr_sources.py:78-87``).  ``evolution`` follows the reference keyword set
(``cr_sources.py:110-162``).  ``BrokenEvolutionSource`` adds the
"(1+z)^m below z_break, flat above" evolution used by the full nuclear
cascade workload; it is a subclass overriding ``evolution`` exactly as the
reference's extension point intends.
"""

from __future__ import annotations

import numpy as np

from .cosmology import star_formation_rate


def _sfr_weighted(z, m):
    return (1 + z) ** m * star_formation_rate(z)


def _broken_at_one(z, m):
    return (1 + z) ** m if z <= 1 else (1 + 1) ** m


def _broken_at_one_to_sfr(z, m):
    return (1 + z) ** m if z <= 1 else (1 + 1) ** 3.6 * (1 + z) ** (m - 3.6)


#: (keyword, m) evolutions of the reference (cr_sources.py:110-162)
EVOLUTIONS = {
    "SFR": _sfr_weighted,
    "TDE": lambda z, m: (1 + z) ** (m - 3.0),
    "simple": lambda z, m: (1 + z) ** m,
    "simple_flat": _broken_at_one,
    "simple_SFR": _broken_at_one_to_sfr,
}


def evolution_factor(m, z):
    """Source density evolution at z for the reference's evolution spec m:
    "flat", a float (then (1+z)^m times the star-formation rate) or
    (keyword, m) with a keyword of ``EVOLUTIONS``."""
    if z < 0:
        raise ValueError(f"source evolution is defined for z >= 0 only (z = {z})")
    if m == "flat":
        return 1.0
    if isinstance(m, float):
        return _sfr_weighted(z, m)
    if isinstance(m, tuple) and len(m) == 2 and m[0] in EVOLUTIONS:
        return EVOLUTIONS[m[0]](z, m[1])
    raise ValueError(f"source evolution {m!r}: expected 'flat', a float or one of "
                     f"{sorted(EVOLUTIONS)} with an index")


class CosmicRaySource:
    """Injection ``norm * evolution(z) * injection_grid`` (cr_sources.py:14-87).

    ``params`` maps PDG ids to the spectrum parameters of the subclass;
    ``injection_grid`` holds every injected species' spectrum in its block
    of the state vector."""

    def __init__(self, prince_run, pdgids=None, params=None, norm=1.0, m="flat"):
        self.spec_man = prince_run.spec_man
        self.cr_grid = prince_run.cr_grid.grid
        self.dim_states = prince_run.dim_states
        self.params = params
        self.norm = norm
        self.source_evo_m = m
        ids = list(params.keys()) if pdgids is None else list(pdgids)
        self.pdgids = np.sort(np.array(ids))
        self.injection_grid = self._injection_grid()

    def _injection_grid(self):
        grid = np.zeros(self.dim_states)
        for pid in (int(p) for p in self.pdgids):
            block = self.spec_man.pdgid2sref[pid].sl
            grid[block] = self.injection_spectrum(pid, self.cr_grid, self.params[pid])
        return grid

    def injection_rate(self, z):
        return self.norm * self.evolution(z) * self.injection_grid

    def injection_spectrum(self, pid, energy, params):  # pragma: no cover
        raise NotImplementedError("subclasses define the injection spectrum")

    def evolution(self, z):
        return evolution_factor(self.source_evo_m, z)

    def injection_coefficients(self, zs):
        """``norm * evolution(z)`` for every z of ``zs`` (the per-step source
        coefficients of a solve).  Point by point with the scalar arithmetic of
        ``evolution`` (same bits); when a subclass keeps this class's
        ``evolution``, the evolution law is looked up once instead of per z."""
        zl = np.asarray(zs, dtype=np.float64).tolist()
        n = self.norm
        if type(self).evolution is CosmicRaySource.evolution:
            m = self.source_evo_m
            if any(z < 0 for z in zl):
                raise ValueError("source evolution is defined for z >= 0 only")
            if m == "flat":
                return np.full(len(zl), n * 1.0)
            if isinstance(m, float):
                return np.array([n * _sfr_weighted(z, m) for z in zl], dtype=np.float64)
            if isinstance(m, tuple) and len(m) == 2 and m[0] in EVOLUTIONS:
                law, mm = EVOLUTIONS[m[0]], m[1]
                return np.array([n * law(z, mm) for z in zl], dtype=np.float64)
        return np.array([n * self.evolution(z) for z in zl], dtype=np.float64)


class SimpleSource(CosmicRaySource):
    """norm E^-gamma exp(-E/Emax), params {pid: (gamma, Emax, norm)}."""

    def injection_spectrum(self, pid, energy, params):
        gamma, emax, norm = params
        inj_spec = self.spec_man.pdgid2sref[pid]
        emax = emax / inj_spec.A
        return norm * energy ** (-gamma) * np.exp(-energy / emax)


class RigidityCutoffSource(CosmicRaySource):
    """params {pid: (gamma, Rcut, norm)}, exp(1 - E_k/(Z Rcut)) everywhere."""

    def injection_spectrum(self, pid, energy, params):
        spectral_index, rcut, relnorm = params
        inj_spec = self.spec_man.pdgid2sref[pid]
        A = float(inj_spec.A)
        emax = rcut * inj_spec.Z
        e_k = A * energy
        return relnorm * A * (e_k / 1e9) ** (-spectral_index) * np.exp(1 - e_k / emax)


class AugerFitSource(CosmicRaySource):
    """Power law with a rigidity cutoff applied above E_k = Z Rcut only."""

    def injection_spectrum(self, pid, energy, params):
        spectral_index, rcut, relnorm = params
        inj_spec = self.spec_man.pdgid2sref[pid]
        A = float(inj_spec.A)
        emax = rcut * inj_spec.Z
        e_k = A * energy
        return (
            relnorm
            * A
            * (e_k / 1e9) ** (-spectral_index)
            * np.where(e_k < emax, 1.0, np.exp(1 - e_k / emax))
        )


class BrokenEvolutionSource(AugerFitSource):
    """AugerFitSource with evolution (1+z)^m for z < z_break, flat above.

    Above the break the evolution stays at its value at the break,
    (1+z_break)^m, so the emissivity is continuous.
    """

    def __init__(self, prince_run, params=None, norm=1.0, m=3.0, z_break=1.5, **kw):
        self.m_break = float(m)
        self.z_break = float(z_break)
        super().__init__(prince_run, params=params, norm=norm, m=("simple", float(m)), **kw)

    def evolution(self, z):
        if z < 0:
            raise ValueError(f"source evolution is defined for z >= 0 only (z = {z})")
        if z < self.z_break:
            return (1 + z) ** self.m_break
        return (1 + self.z_break) ** self.m_break
