"""Named propagation workloads (runs + sources) used by tests and the benchmark.

* ``proton_cmb``: protons only (n folded into p), 1e6-1e14 GeV at 8 bins per
  decade (64 bins), CMB only, SimpleSource E^-2 exp(-E/10^11.5 GeV), no
  evolution, z 1 -> 0.
* ``nuclear_cmb_ebl``: all transported isotopes up to A = 56 (305 species,
  19520 states), CMB + synthetic two-blackbody EBL (seeded amplitudes),
  pure Si-28 with spectral index 1.0 and R_max = 10^9.5 GV, evolution
  (1+z)^3 below z = 1.5 and flat above, z 3 -> 0.
* ``nuclear_batch``: per-instance draws on the nuclear run -- spectral index
  U[-1.5, 2.5], log10 R_max U[8.5, 10.5] GV, composition over
  {H, He, N, Si, Fe} ~ Dirichlet(1, ..., 1), evolution (1+z)^m with
  m ~ U[-6, 6], z_max = 3.
"""

from __future__ import annotations

import numpy as np

from .model.core import PriNCeRun
from .model.cr_sources import AugerFitSource, BrokenEvolutionSource, SimpleSource
from .model.photonfields import CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL
from .model.units import PDG_PROTON, make_nucleus_pdg

SI28 = make_nucleus_pdg(28, 14)
COMPOSITION = (PDG_PROTON, make_nucleus_pdg(4, 2), make_nucleus_pdg(14, 7), SI28,
               make_nucleus_pdg(56, 26))

_RUNS = {}


def proton_run():
    if "proton" not in _RUNS:
        _RUNS["proton"] = PriNCeRun(max_mass=1, photon_field=CMBPhotonSpectrum())
    return _RUNS["proton"]


def nuclear_run(ebl_seed=0):
    key = ("nuclear", ebl_seed)
    if key not in _RUNS:
        pf = CombinedPhotonField([CMBPhotonSpectrum(), SyntheticEBL(seed=ebl_seed)])
        _RUNS[key] = PriNCeRun(max_mass=56, photon_field=pf)
    return _RUNS[key]


def proton_em_run():
    """Protons on the CMB with the co-evolved EM cascade (photons, e+-)."""
    if "proton_em" not in _RUNS:
        _RUNS["proton_em"] = PriNCeRun(max_mass=1, photon_field=CMBPhotonSpectrum(), em_cascade=True)
    return _RUNS["proton_em"]


def nuclear_em_run(ebl_seed=0):
    """The A <= 56 nuclear run (CMB + EBL) with the co-evolved EM cascade."""
    key = ("nuclear_em", ebl_seed)
    if key not in _RUNS:
        pf = CombinedPhotonField([CMBPhotonSpectrum(), SyntheticEBL(seed=ebl_seed)])
        _RUNS[key] = PriNCeRun(max_mass=56, photon_field=pf, em_cascade=True)
    return _RUNS[key]


def proton_source(run):
    return SimpleSource(run, params={PDG_PROTON: (2.0, 10**11.5, 1.0)})


def silicon_source(run):
    return BrokenEvolutionSource(run, params={SI28: (1.0, 10**9.5, 1.0)}, m=3.0, z_break=1.5)


def batch_draws(K, seed=0):
    """Per-instance parameters (gamma, log10 R_max, fractions[5], m)."""
    rng = np.random.default_rng(seed)
    gamma = rng.uniform(-1.5, 2.5, size=K)
    lrmax = rng.uniform(8.5, 10.5, size=K)
    frac = rng.dirichlet(np.ones(len(COMPOSITION)), size=K)
    m = rng.uniform(-6.0, 6.0, size=K)
    return gamma, lrmax, frac, m


def batch_source(run, gamma, lrmax, frac, m):
    params = {pid: (float(gamma), 10 ** float(lrmax), float(f)) for pid, f in zip(COMPOSITION, frac)}
    return AugerFitSource(run, params=params, m=("simple", float(m)))


def batch_sources(run, K, seed=0):
    g, lr, fr, m = batch_draws(K, seed)
    return [batch_source(run, g[k], lr[k], fr[k], m[k]) for k in range(K)]
