"""Units, energy grids, PDG helpers and the species manager.

Host-side model infrastructure used to build the operator objects that the
ETD2 solver consumes.  The conventions follow the reference package:

* units and constants: ``data.py:722-740`` (``PRINCE_UNITS``), cm / s / GeV;
* ``EnergyGrid(lower, upper, bins_dec)``: ``data.py:1068-1090``;
* PDG nucleus codes ``1000000000 + Z*10000 + A*10`` with free p = 2212,
  n = 2112: ``util.py:264-316``;
* ``SpeciesManager``: species sorted by PDG id, one contiguous block of ``d``
  energy bins per species in the state vector: ``data.py:1385-1460``,
  ``1678-1712``.
"""

from __future__ import annotations

from collections import namedtuple

import numpy as np
import scipy.constants as spc

_UNITS = dict(
    c=1e2 * spc.c,
    cm2Mpc=1.0 / (spc.parsec * spc.mega * 1e2),
    Mpc2cm=spc.mega * spc.parsec * 1e2,
    m_proton=spc.physical_constants["proton mass energy equivalent in MeV"][0] * 1e-3,
    m_electron=spc.physical_constants["electron mass energy equivalent in MeV"][0] * 1e-3,
    r_electron=spc.physical_constants["classical electron radius"][0] * 1e2,
    fine_structure=spc.fine_structure,
    GeV2erg=1.0 / 624.15,
    erg2GeV=624.15,
    cm2sec=1e-2 / spc.c,
    sec2cm=spc.c * 1e2,
)

#: Immutable unit/constant record (GeV, cm, s).
PRINCE_UNITS = namedtuple("PriNCeUnits", list(_UNITS.keys()))(**_UNITS)

PDG_PROTON = 2212
PDG_NEUTRON = 2112
_PDG_NUCLEUS_PREFIX = 1000000000


def make_nucleus_pdg(A: int, Z: int) -> int:
    """PDG id of a ground-state nucleus (free p / n get 2212 / 2112)."""
    A, Z = int(A), int(Z)
    if A == 1 and Z == 1:
        return PDG_PROTON
    if A == 1 and Z == 0:
        return PDG_NEUTRON
    return _PDG_NUCLEUS_PREFIX + Z * 10000 + A * 10


def get_AZN(pdg_id: int):
    """(A, Z, N) of a nucleus PDG id; (0, 0, 0) for anything else."""
    pdg_id = int(pdg_id)
    if pdg_id == PDG_PROTON:
        return 1, 1, 0
    if pdg_id == PDG_NEUTRON:
        return 1, 0, 1
    if pdg_id // _PDG_NUCLEUS_PREFIX != 1:
        return 0, 0, 0
    rest = pdg_id - _PDG_NUCLEUS_PREFIX
    Z = (rest // 10000) % 1000
    A = (rest // 10) % 1000
    return A, Z, A - Z


def is_nucleus(pdg_id: int) -> bool:
    A, Z, _ = get_AZN(pdg_id)
    return A >= 1 and Z <= A


class EnergyGrid:
    """Logarithmic energy grid (reference ``data.py:1068-1090``).

    ``bins`` are the edges ``logspace(lower, upper, (upper-lower)*bins_dec+1)``,
    ``grid`` the arithmetic bin centres, ``widths`` the bin widths.
    """

    def __init__(self, lower, upper, bins_dec):
        self.bins = np.logspace(lower, upper, int((upper - lower) * bins_dec + 1))
        self.grid = 0.5 * (self.bins[1:] + self.bins[:-1])
        self.widths = self.bins[1:] - self.bins[:-1]
        self.d = self.grid.size


class PrinceSpecies:
    """One transported species: PDG id, (A, Z, N), charge and its block."""

    def __init__(self, pdgid: int, princeidx: int, d: int):
        self.pdgid = int(pdgid)
        self.princeidx = int(princeidx)
        self.is_nucleus = is_nucleus(pdgid)
        self.A, self.Z, self.N = get_AZN(pdgid) if self.is_nucleus else (0, 0, 0)
        self.charge = self.Z
        self.grid_tag = "default"
        self._tr_dim = d
        self._tr_lidx = princeidx * d
        self._tr_lbin = princeidx * (d + 1)

    @property
    def sl(self):
        return slice(self._tr_lidx, self._tr_lidx + self._tr_dim)

    def lidx(self):
        return self._tr_lidx

    def uidx(self):
        return self._tr_lidx + self._tr_dim

    def lbin(self):
        return self._tr_lbin

    def ubin(self):
        return self._tr_lbin + self._tr_dim + 1

    def __repr__(self):
        return f"PrinceSpecies({self.pdgid}, A={self.A}, Z={self.Z}, idx={self.princeidx})"


class SpeciesManager:
    """Species table: sorted unique PDG ids, block ``princeidx*d`` each."""

    def __init__(self, pdgid_list, ed: int):
        pdgid_list = sorted(set(int(p) for p in pdgid_list))
        self.grid_dims = {"default": ed}
        self.species_refs = [PrinceSpecies(p, i, ed) for i, p in enumerate(pdgid_list)]
        self.known_species = [s.pdgid for s in self.species_refs]
        self.pdgid2sref = {s.pdgid: s for s in self.species_refs}
        self.pdgid2princeidx = {s.pdgid: s.princeidx for s in self.species_refs}
        self.princeidx2sref = {s.princeidx: s for s in self.species_refs}
        self.princeidx2pdgid = {s.princeidx: s.pdgid for s in self.species_refs}
        self.nspec = len(self.species_refs)
        self.dim_states = ed * self.nspec
        self.dim_bins = (ed + 1) * self.nspec

    def add_grid(self, grid_tag, dimension):
        self.grid_dims[grid_tag] = dimension
