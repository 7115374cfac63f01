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


class CosmicRaySource:
    def __init__(self, prince_run, pdgids=None, params=None, norm=1.0, m="flat"):
        self.cr_grid = prince_run.cr_grid.grid
        self.dim_states = prince_run.dim_states
        self.norm = norm
        self.spec_man = prince_run.spec_man
        self.params = params
        self.pdgids = np.array(pdgids if pdgids is not None else list(params.keys()))
        self.pdgids.sort()
        self.source_evo_m = m
        self._compute_injection_grid()

    def _compute_injection_grid(self):
        self.injection_grid = np.zeros(self.dim_states)
        for pid in self.pdgids:
            params = self.params[pid]
            inj_spec = self.spec_man.pdgid2sref[int(pid)]
            self.injection_grid[inj_spec.sl] = self.injection_spectrum(
                int(pid), self.cr_grid, params
            )

    def injection_rate(self, z):
        return self.norm * self.evolution(z) * self.injection_grid

    def injection_spectrum(self, pid, energy, params):  # pragma: no cover
        raise NotImplementedError

    def evolution(self, z):
        if z < 0:
            raise Exception("Source evolution not defined for negative z = {:}".format(z))
        m = self.source_evo_m
        if m == "flat":
            return 1.0
        elif type(m) is float:
            return (1 + z) ** m * star_formation_rate(z)
        elif type(m) is tuple:
            mkwd, mval = m
            if mkwd == "SFR":
                return (1 + z) ** mval * star_formation_rate(z)
            elif mkwd == "TDE":
                return (1 + z) ** (mval - 3.0)
            elif mkwd == "simple":
                return (1 + z) ** mval
            elif mkwd == "simple_flat":
                if z <= 1:
                    return (1 + z) ** mval
                return (1 + 1) ** mval
            elif mkwd == "simple_SFR":
                if z <= 1:
                    return (1 + z) ** mval
                return (1 + 1) ** 3.6 * (1 + z) ** (mval - 3.6)
        raise Exception("Unknown source evo type: {:}".format(m))


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
            raise Exception("Source evolution not defined for negative z = {:}".format(z))
        if z < self.z_break:
            return (1 + z) ** self.m_break
        return (1 + self.z_break) ** self.m_break
