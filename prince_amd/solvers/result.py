"""Propagation result container (reference ``solvers/propagation.py:34-192``).

Holds the final state vector on the cosmic-ray grid and exposes the same
accessors as the reference: per-species spectra per nucleon
(``get_solution``), in total energy with an energy power
(``get_solution_scale``), group sums and the mean / variance of ln A on a
common total-energy grid obtained by log-log interpolation
(``get_solution_group``, ``get_lnA``), plus ``to_dict`` / ``from_dict``.
"""

from __future__ import annotations

import numpy as np

from ..model import config
from ..model.units import EnergyGrid, SpeciesManager, is_nucleus


class UHECRPropagationResult:
    """``em`` maps the PDG ids of EM species (photon, e+-) to (grid, slice of
    state) when the solve carried the EM cascade; they live on their own grid."""

    def __init__(self, state, egrid, spec_man, em=None):
        self.spec_man = spec_man
        self.egrid = egrid
        self.state = state
        self.em = dict(em or {})

    def to_dict(self):
        return {"egrid": self.egrid, "state": self.state, "known_spec": self.known_species}

    @classmethod
    def from_dict(cls, dic):
        egrid = dic["egrid"]
        spec_man = SpeciesManager(dic["known_spec"], egrid.size)
        return cls(dic["state"], egrid, spec_man)

    @property
    def known_species(self):
        return self.spec_man.known_species

    def __add__(self, other):
        if not np.array_equal(self.egrid, other.egrid):
            raise Exception(
                "Cannot add Propagation Results, they are defined in different energy grids!"
            )
        if not np.array_equal(self.known_species, other.known_species):
            raise Exception("Cannot add Propagation Results, have different species managers!")
        return UHECRPropagationResult(self.state + other.state, self.egrid, self.spec_man, em=self.em)

    def __mul__(self, number):
        if not np.isscalar(number):
            raise Exception(
                "Can only multiply result by scalar number, got type {:} instead!".format(
                    type(number)
                )
            )
        return UHECRPropagationResult(self.state * number, self.egrid, self.spec_man, em=self.em)

    def get_solution(self, pdg_id):
        if pdg_id in self.em:
            grid, sl = self.em[pdg_id]
            return grid, self.state[sl]
        spec = self.spec_man.pdgid2sref[pdg_id]
        return self.egrid, self.state[spec.lidx() : spec.uidx()]

    def get_solution_scale(self, pdg_id, epow=0):
        spec = self.spec_man.pdgid2sref[pdg_id]
        egrid = spec.A * self.egrid
        return egrid, egrid**epow * self.state[spec.lidx() : spec.uidx()] / spec.A

    def _check_id_grid(self, pdg_ids, egrid):
        if egrid is None:
            max_mass = max(s.A for s in self.spec_man.species_refs)
            emin_log, emax_log, nbins = list(config.cosmic_ray_grid)
            emax_log = np.log10(max_mass) + emax_log
            nbins *= 4
            com_egrid = EnergyGrid(emin_log, emax_log, nbins).grid
        else:
            com_egrid = egrid
        if isinstance(pdg_ids, list):
            pass
        elif pdg_ids == "CR":
            pdg_ids = [s for s in self.known_species if is_nucleus(s)]
        elif pdg_ids == "all":
            pdg_ids = self.known_species
        elif isinstance(pdg_ids, tuple):
            select, vmin, vmax = pdg_ids
            pdg_ids = [s for s in self.known_species if vmin <= select(s) <= vmax]
        return pdg_ids, com_egrid

    def _collect_interpolated_spectra(self, pdg_ids, epow, egrid=None):
        pdg_ids, com_egrid = self._check_id_grid(pdg_ids, egrid)
        spectra = np.zeros((len(pdg_ids), com_egrid.size))
        for idx, pid in enumerate(pdg_ids):
            curr_egrid, curr_spec = self.get_solution_scale(pid, epow)
            mask = curr_spec > 0.0
            if np.count_nonzero(mask) > 0:
                res = np.exp(
                    np.interp(
                        np.log(com_egrid),
                        np.log(curr_egrid[mask]),
                        np.log(curr_spec[mask]),
                        left=np.nan,
                        right=np.nan,
                    )
                )
            else:
                res = np.zeros_like(com_egrid)
            spectra[idx] = np.nan_to_num(res)
        return pdg_ids, com_egrid, spectra

    def get_solution_group(self, pdg_ids, epow=3, egrid=None):
        _, com_egrid, spectra = self._collect_interpolated_spectra(pdg_ids, epow, egrid)
        return com_egrid, spectra.sum(axis=0)

    def get_lnA(self, pdg_ids, egrid=None):
        pdg_ids, com_egrid, spectra = self._collect_interpolated_spectra(pdg_ids, 0, egrid)
        lnA = np.array([np.log(self.spec_man.pdgid2sref[el].A) for el in pdg_ids])
        total = spectra.sum(axis=0)
        with np.errstate(invalid="ignore", divide="ignore"):
            average = (lnA[:, np.newaxis] * spectra).sum(axis=0) / total
            variance = (lnA[:, np.newaxis] ** 2 * spectra).sum(axis=0) / total - average**2
        return com_egrid, average, variance

    def get_energy_density(self, pdg_id):
        from scipy.integrate import trapezoid

        A = self.spec_man.pdgid2sref[pdg_id].A
        return trapezoid(A * self.egrid * self.get_solution(pdg_id)[1], self.egrid)
