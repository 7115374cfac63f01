"""Synthetic photo-nuclear cross-section model and its response functions.

The reference loads tabulated cross sections from external FLUKA/SOPHIA
databases (``cross_sections/fluka.py:171-336``, ``photo_meson.py:43-89``),
which do not exist offline.  This module builds tables of the same kinds on
the same kinds of grids, so the rate-kernel construction that consumes them
(``interaction_rates.py:251-646``) runs unchanged:

* ``nonel(mo)``: non-elastic cross section sigma(eps_r) on a shared photon
  energy grid ``egrid`` (nucleus rest frame, GeV, cm^2);
* ``incl(mo, da)``: boost-conserving inclusive cross sections for nuclear
  daughters (A >= 2);
* ``incl_diff(mo, da)``: d sigma/dx(x, eps_r) for redistributed daughters
  (free nucleons), on ``xcenters`` of log-spaced ``xbins`` in [1e-5, 3]
  (reference ``base.py:556-652``).

Physics content (a model, not data): giant dipole resonance (Lorentzian with
Berman-Fultz centroid and a TRK-sum-rule strength), a quasi-deuteron plateau
and a Delta(1232) + multi-pion photo-pion part scaling with A.  GDR strength
is split into 1n, 1p, 2n, pn and alpha emission; photo-pion strength into
spallation fragments A-1..A-8.  Daughters that are not in the species set are
mapped to the isotope of the same mass with the nearest charge (what the
reference's decay-chain reducer does for beta-unstable daughters, which keeps
A), or dropped when no isotope of that mass is transported.  Emitted
nucleons (neutrons decay into protons, as with ``tau_dec_threshold=inf``)
feed one redistributed (mo -> p) channel: a narrow component at x ~ 1 from
evaporation and a broad photo-pion component.

``ResponseFunction`` restates ``cross_sections/response.py:99-274``:
f(y) = cumtrapz(eps sigma, eps) / (2 y^2) on y = egrid[1:]/2, and for
differential channels the 2-D double antiderivative over (x, y).
"""

from __future__ import annotations

import numpy as np
from scipy.integrate import cumulative_trapezoid

from .units import PDG_PROTON, get_AZN, make_nucleus_pdg

MB = 1e-27  # cm^2


def _z_stable(A):
    return A / (1.98 + 0.0155 * A ** (2.0 / 3.0))


def nuclear_species(max_mass=56, width=(3, 2), extra=()):
    """Isotopes A <= max_mass in a charge band around the stability line.

    ``width = (below, above)`` keeps Z in [Z_s - below, Z_s + above] where
    Z_s = A / (1.98 + 0.0155 A^(2/3)) rounded; an int means symmetric.

    A = 5 and A = 8 are left out (no bound isotopes).  Always contains the
    proton and the injection nuclei He-4, N-14, Si-28, Fe-56 when in range.
    """
    out = {PDG_PROTON}
    for A in range(2, int(max_mass) + 1):
        if A in (5, 8):
            continue
        zs = int(round(_z_stable(A)))
        wlo, whi = (width, width) if isinstance(width, int) else width
        for Z in range(zs - wlo, zs + whi + 1):
            if Z < 1 or Z > A or (A - Z) < 1 and A > 1:
                continue
            if A >= 2 and (A - Z) < 0:
                continue
            # no bare multi-proton or multi-neutron systems
            if A >= 3 and (Z == A or Z == 0):
                continue
            out.add(make_nucleus_pdg(A, Z))
    for A, Z in ((4, 2), (14, 7), (28, 14), (56, 26)):
        if A <= max_mass:
            out.add(make_nucleus_pdg(A, Z))
    for p in extra:
        out.add(int(p))
    return sorted(out)


def _smooth_threshold(eps, eth, power=1.5):
    r = np.clip(1.0 - eth / np.maximum(eps, 1e-300), 0.0, None)
    return r**power


def sigma_photopion_proton(eps):
    """Photo-pion cross section of a free nucleon (cm^2) vs rest-frame eps."""
    eth = 0.145
    delta = 0.48 * MB * (0.11**2 * eps**2) / ((eps**2 - 0.34**2) ** 2 + 0.11**2 * eps**2)
    multi = 0.12 * MB * _smooth_threshold(eps, 0.5, 2.0)
    return (delta * (eps / 0.34) ** 0 + multi) * _smooth_threshold(eps, eth, 1.0)


def sigma_gdr(eps, A, Z):
    N = A - Z
    e0 = (31.2 * A ** (-1.0 / 3.0) + 20.6 * A ** (-1.0 / 6.0)) * 1e-3
    gam = (10.0 - 5.0 * min(A, 56) / 56.0) * 1e-3
    strength = 60.0 * N * Z / A * 1e-3 * MB  # integral sigma d eps (GeV cm^2)
    s0 = 2.0 * strength / (np.pi * gam)
    lor = s0 * (eps**2 * gam**2) / ((eps**2 - e0**2) ** 2 + eps**2 * gam**2)
    eth = 2.2e-3 if A == 2 else (6.0e-3 if A <= 4 else 8.0e-3)
    qd = 0.25 * N * Z / A * MB * _smooth_threshold(eps, 0.03, 1.0) * np.exp(-eps / 0.15)
    return (lor + qd) * _smooth_threshold(eps, eth, 1.5)


class SyntheticPhotoNuclear:
    """Channel catalogue plus tables on shared (egrid, xbins) grids."""

    GDR_SPLIT = ((1, 0, 0.60), (1, 1, 0.25), (2, 0, 0.07), (2, 1, 0.05), (4, 2, 0.03))
    PION_SPLIT = tuple((k, 0.5 * 0.6 ** (k - 1)) for k in range(1, 9))

    def __init__(self, species, egrid=None, nx=100):
        self.species = sorted(int(s) for s in species)
        if egrid is None:
            egrid = np.logspace(-3, 10, 13 * 24 + 1)
        self._egrid = np.asarray(egrid, dtype=np.float64)
        self.xbins = np.logspace(-5, np.log10(3.0), nx + 1)
        self._by_A = {}
        for s in self.species:
            A, Z, _ = get_AZN(s)
            self._by_A.setdefault(A, []).append(Z)
        self._build_tables()
        self._resp = None

    # ------------------------------------------------------------------
    @property
    def egrid(self):
        return self._egrid

    @property
    def xcenters(self):
        return 0.5 * (self.xbins[1:] + self.xbins[:-1])

    @property
    def xwidths(self):
        return self.xbins[1:] - self.xbins[:-1]

    def _map_daughter(self, A, Z):
        """Transported isotope of mass A closest in charge to Z (None if none)."""
        if A < 2:
            return None
        zs = self._by_A.get(A)
        if not zs:
            return None
        zbest = min(zs, key=lambda zz: (abs(zz - Z), -zz))
        return make_nucleus_pdg(A, zbest)

    def _x_profiles(self):
        xc = self.xcenters
        narrow = np.exp(-0.5 * (np.log(xc) / 0.12) ** 2)
        broad = xc**0.3 * np.exp(-xc / 0.45)
        narrow /= np.trapezoid(narrow, xc)
        broad /= np.trapezoid(broad, xc)
        leading = np.exp(-0.5 * ((xc - 0.8) / 0.12) ** 2) * (xc < 1.2)
        leading /= np.trapezoid(leading, xc)
        return narrow, broad, leading

    def _build_tables(self):
        eps = self._egrid
        narrow, broad, leading = self._x_profiles()
        self._nonel_tab = {}
        self._incl_tab = {}
        self._incl_diff_tab = {}
        sig_pi_N = sigma_photopion_proton(eps)
        for mo in self.species:
            A, Z, N = get_AZN(mo)
            if mo == PDG_PROTON:
                self._nonel_tab[mo] = sig_pi_N.copy()
                # leading nucleon (p or n -> p): multiplicity 1, x near 0.8
                self._incl_diff_tab[(mo, PDG_PROTON)] = np.outer(leading, sig_pi_N)
                continue
            gdr = sigma_gdr(eps, A, Z)
            pion = 0.9 * A * sig_pi_N
            self._nonel_tab[mo] = gdr + pion
            nucleon_mult = np.zeros_like(eps)
            # GDR emission channels
            he4 = make_nucleus_pdg(4, 2)
            have_he4 = 2 in self._by_A.get(4, [])
            for dA, dZ, frac in self.GDR_SPLIT:
                Ad = A - dA
                if Ad < 1:
                    continue
                w = frac * gdr * (1.0 if dA == 1 else _smooth_threshold(eps, 0.02, 1.0))
                if Ad == 1:
                    # break-up into free nucleons only
                    nucleon_mult = nucleon_mult + A * w
                    continue
                da = self._map_daughter(Ad, Z - dZ)
                if da is not None:
                    self._incl_tab[(mo, da)] = self._incl_tab.get((mo, da), 0.0) + w
                if dA == 4:
                    if have_he4 and he4 != mo:
                        self._incl_tab[(mo, he4)] = self._incl_tab.get((mo, he4), 0.0) + w
                else:
                    nucleon_mult = nucleon_mult + dA * w
            # photo-pion spallation fragments A-k
            for k, frac in self.PION_SPLIT:
                Ad = A - k
                if Ad < 2:
                    continue
                da = self._map_daughter(Ad, int(round(Z * Ad / A)))
                if da is None:
                    continue
                w = frac * pion
                self._incl_tab[(mo, da)] = self._incl_tab.get((mo, da), 0.0) + w
                nucleon_mult = nucleon_mult + k * w
            # emitted nucleons: evaporation near x=1, knock-on broad
            self._incl_diff_tab[(mo, PDG_PROTON)] = np.outer(narrow, nucleon_mult) + np.outer(
                broad, 0.5 * pion
            )
        # drop self-couplings and empty channels
        self._incl_tab = {
            k: v for k, v in self._incl_tab.items() if k[0] != k[1] and np.any(v > 0)
        }
        self.nonel_idcs = sorted(self._nonel_tab)
        self.incl_idcs = sorted(self._incl_tab)
        self.incl_diff_idcs = sorted(self._incl_diff_tab)
        self.known_species = list(self.species)
        self.known_bc_channels = list(self.incl_idcs)
        self.known_diff_channels = list(self.incl_diff_idcs)
        self.known_bc_channels_set = set(self.known_bc_channels)
        self.known_diff_channels_set = set(self.known_diff_channels)
        self.reactions = {mo: [] for mo in self.species}
        for key in self.incl_idcs + self.incl_diff_idcs:
            self.reactions[key[0]].append(key)

    # ------------------------------------------------------------------
    def nonel(self, mother):
        return self._egrid, self._nonel_tab[mother]

    def incl(self, mother, daughter):
        if (mother, daughter) in self._incl_diff_tab:
            cs = np.trapezoid(self._incl_diff_tab[(mother, daughter)], x=self.xcenters, axis=0)
            return self._egrid, cs
        return self._egrid, self._incl_tab[(mother, daughter)]

    def incl_diff(self, mother, daughter):
        return self._egrid, self._incl_diff_tab[(mother, daughter)]

    @property
    def resp(self):
        if self._resp is None:
            self._resp = ResponseFunction(self)
        return self._resp


class ResponseFunction:
    """Response functions of all channels on the shared y grid.

    Stores the 1-D stacks ``nonel_f_stack`` / ``incl_f_stack`` (rows aligned
    with ``nonel_keys`` / ``incl_keys``) of f(y) = cumtrapz(eps sigma)/(2y^2)
    on ``response_ygrid = egrid[1:]/2`` (response.py:163-205) and, for every
    differential channel, the double antiderivative
    ``int_0^x int_0^y h`` on (xcenters, ygrid) (response.py:220-245),
    stacked along the last axis in ``incl_diff_integral_keys`` order.
    """

    def __init__(self, cs: SyntheticPhotoNuclear):
        self.cross_section = cs
        egrid = cs.egrid
        ygr = egrid[1:] / 2.0
        self.response_ygrid = ygr
        self.xcenters = cs.xcenters
        self.xbins = cs.xbins

        def _stack(keys, getter):
            if not keys:
                return np.zeros((0, ygr.size))
            sig = np.array([getter(k)[1] for k in keys])
            return cumulative_trapezoid(egrid * sig, x=egrid, axis=1) / (2.0 * ygr**2)

        self.nonel_keys = list(cs.nonel_idcs)
        self.incl_keys = list(cs.incl_idcs)
        self.nonel_f_stack = _stack(self.nonel_keys, lambda mo: cs.nonel(mo))
        self.incl_f_stack = _stack(self.incl_keys, lambda k: cs.incl(*k))

        self.incl_diff_integral_keys = list(cs.incl_diff_idcs)
        integrals = []
        for key in self.incl_diff_integral_keys:
            eg, sig2 = cs.incl_diff(*key)
            rfunc = cumulative_trapezoid(eg * sig2, x=eg, axis=1) / (2.0 * ygr**2)
            integral = cumulative_trapezoid(rfunc, ygr, axis=1, initial=0)
            integral = cumulative_trapezoid(integral, self.xcenters, axis=0, initial=0)
            integrals.append(integral)
        self.incl_diff_integral_stack = (
            np.stack(integrals, axis=-1) if integrals else np.zeros((cs.xcenters.size, ygr.size, 0))
        )
