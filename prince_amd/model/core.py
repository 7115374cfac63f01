"""Run container: grids, species, photon field, loss rates and rate kernel.

Mirrors the construction order of the reference's ``PriNCeRun``
(``core.py:19-257``): energy grids -> cross sections -> photon field ->
species manager (sorted PDG ids) -> continuous loss rates -> photo-nuclear
interaction rates.  The cross sections default to the synthetic model of
``xsec.py``; the photon field defaults to the CMB.
"""

from __future__ import annotations

from . import cascade, config
from .interaction_rates import (
    ContinuousAdiabaticLossRate,
    ContinuousPairProductionLossRate,
    PhotoNuclearInteractionRate,
)
from .photonfields import CMBPhotonSpectrum
from .units import EnergyGrid, SpeciesManager, get_AZN
from .xsec import SyntheticPhotoNuclear, nuclear_species


#: EM species in PDG order: e+, e-, photon
EM_PDGS = (-11, 11, 22)


class PriNCeRun:
    def __init__(
        self,
        max_mass=56,
        species_list=None,
        photon_field=None,
        cross_sections=None,
        cosmic_ray_grid=None,
        photon_grid=None,
        backend=None,
        em_cascade=False,
        kernel_cache=None,
    ):
        self.backend = backend if backend is not None else config.BackendConfig()
        self.cr_grid = EnergyGrid(*(cosmic_ray_grid or config.cosmic_ray_grid))
        self.ph_grid = EnergyGrid(*(photon_grid or config.photon_grid))
        if species_list is None:
            species_list = [s for s in nuclear_species(max_mass) if get_AZN(s)[0] <= max_mass]
        if cross_sections is None:
            cross_sections = SyntheticPhotoNuclear(species_list)
        self.cross_sections = cross_sections
        self.photon_field = photon_field if photon_field is not None else CMBPhotonSpectrum()
        system_species = [s for s in species_list if s in set(cross_sections.known_species)]
        self.spec_man = SpeciesManager(system_species, self.cr_grid.d)
        self.dim_states = self.spec_man.dim_states
        self.dim_bins = self.spec_man.dim_bins
        self.adia_loss_rates_grid = ContinuousAdiabaticLossRate(self, energy="grid")
        self.pair_loss_rates_grid = ContinuousPairProductionLossRate(self, energy="grid")
        self.adia_loss_rates_bins = ContinuousAdiabaticLossRate(self, energy="bins")
        self.pair_loss_rates_bins = ContinuousPairProductionLossRate(self, energy="bins")
        self.int_rates = PhotoNuclearInteractionRate(self, kernel_cache=kernel_cache)
        self.em_int_rates = None
        self.spec_man.add_grid("ph", self.ph_grid.d)
        # co-evolved EM cascade (propagation.py:250-316): photons and e+- on
        # their own 136-bin grid, driven by Bethe-Heitler pairs of the nuclei
        self.enable_em_cascade = bool(em_cascade)
        self.em_grid = EnergyGrid(*cascade.EM_GRID) if em_cascade else None
        self.em_pdgs = EM_PDGS if em_cascade else ()

    @property
    def dim_em(self):
        """Length of the EM part of the state (e+, e-, photon blocks)."""
        return len(self.em_pdgs) * (self.em_grid.d if self.em_grid is not None else 0)

    def em_slices(self):
        """{pdg: slice} of the EM blocks, placed after the nuclear states."""
        if not self.enable_em_cascade:
            return {}
        d = self.em_grid.d
        return {p: slice(self.dim_states + i * d, self.dim_states + (i + 1) * d)
                for i, p in enumerate(self.em_pdgs)}

    def set_photon_field(self, pfield):
        self.photon_field = pfield
