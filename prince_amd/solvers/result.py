"""Final fluxes of a propagation solve.

Same interface as the reference's ``UHECRPropagationResult``
(``solvers/propagation.py:34-192``): ``get_solution(pdg)`` (per-nucleon
spectrum on the solver grid), ``get_solution_scale(pdg, epow)`` (total-energy
grid, E^epow weighted, per nucleus), ``get_solution_group(ids, epow, egrid)``
and ``get_lnA(ids, egrid)`` on a common total-energy grid, ``+`` and scalar
``*``, ``to_dict`` / ``from_dict`` (keys ``egrid``, ``state``,
``known_spec``).

Common grid: unless given, 1e6 .. A_max * 1e14 GeV at four times the solver's
bins per decade.  A species spectrum reaches it by linear interpolation of
log(flux) in log(E) through the species' positive bins only; outside the
span of those bins it contributes zero.

``ResultBatch`` does the same for the K columns of a multi-RHS state at once
(one interpolation plan per species, applied to every column).
"""

from __future__ import annotations

import numpy as np
from scipy.integrate import trapezoid

from ..model import config
from ..model.units import EnergyGrid, SpeciesManager, is_nucleus


def default_common_grid(spec_man):
    """Total-energy grid shared by all species of a result."""
    a_max = max(s.A for s in spec_man.species_refs)
    lo, hi, per_decade = config.cosmic_ray_grid
    return EnergyGrid(lo, np.log10(a_max) + hi, 4 * per_decade).grid


def select_species(spec_man, pdg_ids):
    """PDG ids named by a selector: a list, ``"CR"`` (nuclei), ``"all"``, or
    ``(key, vmin, vmax)`` -- ids with vmin <= key(id) <= vmax."""
    known = list(spec_man.known_species)
    if isinstance(pdg_ids, (list, np.ndarray)):
        return list(pdg_ids)
    if isinstance(pdg_ids, tuple):
        key, vmin, vmax = pdg_ids
        return [p for p in known if vmin <= key(p) <= vmax]
    if pdg_ids == "CR":
        return [p for p in known if is_nucleus(p)]
    if pdg_ids == "all":
        return known
    raise ValueError(f"unknown species selector {pdg_ids!r}")


def loglog_onto(grid_to, grid_from, values):
    """values(grid_from) -> grid_to by linear interpolation of log v in log E
    through the positive samples; zero outside their span (and if none)."""
    pos = values > 0.0
    if not pos.any():
        return np.zeros(grid_to.size)
    lx = np.log(grid_from[pos])
    lt = np.log(grid_to)
    inside = (lt >= lx[0]) & (lt <= lx[-1])
    out = np.zeros(grid_to.size)
    out[inside] = np.exp(np.interp(lt[inside], lx, np.log(values[pos])))
    return out


class UHECRPropagationResult:
    """``em`` maps the PDG ids of EM species (photon, e+-) to (grid, slice of
    state) when the solve carried the EM cascade; they live on their own grid."""

    def __init__(self, state, egrid, spec_man, em=None):
        self.spec_man = spec_man
        self.egrid = egrid
        self.state = state
        self.em = dict(em or {})

    # ------------------------------------------------------------ containers
    @property
    def known_species(self):
        return self.spec_man.known_species

    def to_dict(self):
        return {"egrid": self.egrid, "state": self.state, "known_spec": self.known_species}

    @classmethod
    def from_dict(cls, dic):
        grid = np.asarray(dic["egrid"])
        return cls(dic["state"], grid, SpeciesManager(dic["known_spec"], grid.size))

    def _same_layout(self, other):
        if not np.array_equal(self.egrid, other.egrid):
            raise ValueError("results on different energy grids cannot be added")
        if not np.array_equal(self.known_species, other.known_species):
            raise ValueError("results with different species lists cannot be added")

    def __add__(self, other):
        self._same_layout(other)
        return type(self)(self.state + other.state, self.egrid, self.spec_man, em=self.em)

    def __mul__(self, number):
        if not np.isscalar(number):
            raise TypeError(f"a result scales by a scalar only, not by {type(number).__name__}")
        return type(self)(self.state * number, self.egrid, self.spec_man, em=self.em)

    # ------------------------------------------------------------ spectra
    def _block(self, pdg_id):
        s = self.spec_man.pdgid2sref[pdg_id]
        return s, self.state[s.lidx():s.uidx()]

    def get_solution(self, pdg_id):
        """(grid, flux per nucleon energy) of one species on its own grid."""
        if pdg_id in self.em:
            grid, sl = self.em[pdg_id]
            return grid, self.state[sl]
        _, block = self._block(pdg_id)
        return self.egrid, block

    def get_solution_scale(self, pdg_id, epow=0):
        """(total-energy grid A E, (A E)^epow * flux / A) of one nuclear species."""
        s, block = self._block(pdg_id)
        e_tot = s.A * self.egrid
        return e_tot, e_tot**epow * block / s.A

    def _check_id_grid(self, pdg_ids, egrid):
        grid = default_common_grid(self.spec_man) if egrid is None else egrid
        return select_species(self.spec_man, pdg_ids), grid

    def _collect_interpolated_spectra(self, pdg_ids, epow, egrid=None):
        ids, grid = self._check_id_grid(pdg_ids, egrid)
        spectra = np.zeros((len(ids), grid.size))
        for row, pid in enumerate(ids):
            e_tot, flux = self.get_solution_scale(pid, epow)
            spectra[row] = loglog_onto(grid, e_tot, flux)
        return ids, grid, spectra

    def get_solution_group(self, pdg_ids, epow=3, egrid=None):
        _, grid, spectra = self._collect_interpolated_spectra(pdg_ids, epow, egrid)
        return grid, spectra.sum(axis=0)

    def get_lnA(self, pdg_ids, egrid=None):
        """(grid, <ln A>, Var(ln A)) weighted by the species fluxes (epow 0)."""
        ids, grid, spectra = self._collect_interpolated_spectra(pdg_ids, 0, egrid)
        lnA = np.log([self.spec_man.pdgid2sref[p].A for p in ids])[:, None]
        total = spectra.sum(axis=0)
        with np.errstate(invalid="ignore", divide="ignore"):
            mean = (lnA * spectra).sum(axis=0) / total
            second = (lnA**2 * spectra).sum(axis=0) / total
        return grid, mean, second - mean**2

    def get_energy_density(self, pdg_id):
        s, block = self._block(pdg_id)
        return trapezoid(s.A * self.egrid * block, self.egrid)


class ResultBatch:
    """Observables of the K columns of a multi-RHS state [dim_states, K] at once.

    ``group(ids, epow)`` and ``lnA(ids)`` give [grid, K] arrays equal, column
    by column, to ``UHECRPropagationResult(state[:, k], ...)``'s
    ``get_solution_group`` / ``get_lnA``.  Per species the positive knots of
    every column are found with cumulative max/min scans, and the log-log
    interpolation of all columns is one gather.
    """

    def __init__(self, state, egrid, spec_man, egrid_common=None):
        self.state = np.asarray(state, dtype=np.float64)
        if self.state.ndim != 2:
            raise ValueError("ResultBatch takes a [dim_states, K] state")
        self.egrid = np.asarray(egrid)
        self.spec_man = spec_man
        self.grid = default_common_grid(spec_man) if egrid_common is None else egrid_common

    @property
    def K(self):
        return self.state.shape[1]

    def _species_on_grid(self, pid, epow):
        """[grid, K] interpolated (A E)^epow flux / A of one species for every column."""
        s = self.spec_man.pdgid2sref[pid]
        d = self.egrid.size
        block = self.state[s.lidx():s.uidx()]  # [d, K]
        e_tot = s.A * self.egrid
        vals = (e_tot**epow)[:, None] * block / s.A
        pos = vals > 0.0
        idx = np.arange(d)[:, None]
        prev = np.maximum.accumulate(np.where(pos, idx, -1), axis=0)  # last positive <= i
        nxt = np.minimum.accumulate(np.where(pos, idx, d)[::-1], axis=0)[::-1]  # first >= i
        lt = np.log(self.grid)
        le = np.log(e_tot)
        b = np.searchsorted(le, lt, side="right") - 1  # last bin with le <= lt (-1: none)
        c = np.searchsorted(le, lt, side="left")  # first bin with le >= lt (d: none)
        lo = np.where((b >= 0)[:, None], prev[np.clip(b, 0, d - 1)], -1)  # [n_grid, K]
        hi = np.where((c < d)[:, None], nxt[np.clip(c, 0, d - 1)], d)
        ok = (lo >= 0) & (hi < d)
        lo_c, hi_c = np.where(ok, lo, 0), np.where(ok, hi, 0)
        cols = np.arange(self.K)[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            v0 = np.log(np.where(ok, vals[lo_c, cols], 1.0))
            v1 = np.log(np.where(ok, vals[hi_c, cols], 1.0))
            x0, x1 = le[lo_c], le[hi_c]
            between = hi_c > lo_c
            t = np.where(between, (lt[:, None] - x0) / np.where(between, x1 - x0, 1.0), 0.0)
            out = np.exp(v0 + t * (v1 - v0))
        return np.where(ok, out, 0.0)

    def group(self, pdg_ids="all", epow=3):
        ids = select_species(self.spec_man, pdg_ids)
        tot = np.zeros((self.grid.size, self.K))
        for p in ids:
            tot += self._species_on_grid(p, epow)
        return self.grid, tot

    def lnA(self, pdg_ids="all"):
        ids = select_species(self.spec_man, pdg_ids)
        tot = np.zeros((self.grid.size, self.K))
        m1 = np.zeros_like(tot)
        m2 = np.zeros_like(tot)
        for p in ids:
            sp = self._species_on_grid(p, 0)
            la = np.log(self.spec_man.pdgid2sref[p].A)
            tot += sp
            m1 += la * sp
            m2 += la * la * sp
        with np.errstate(invalid="ignore", divide="ignore"):
            mean = m1 / tot
            return self.grid, mean, m2 / tot - mean**2
