"""Electromagnetic cascade operators: the per-window transfer T(z) and the
Bethe-Heitler pair shape R_BH(z).

Physics (reference ``cascade/kernels.py``, ``cascade/cascade.py``,
``cascade/opacity.py``, ``cascade/bethe_heitler.py``), GeV / cm / s units:

* inverse Compton emission of an electron of Lorentz factor g on the target
  field n(eps): the Blumenthal & Gould (1970) kernel
  ``(3/4) sigma_T c / g^2 * integral n(eps)/eps f(q, Gamma) d eps`` with
  ``Gamma = 4 eps g / m_e``, ``q = E1 / (Gamma (E_e - E1))``,
  ``1/(4 g^2) <= q <= 1`` (``kernels.py:68-106``, cached tensor form
  ``cascade.py:78-104``);
* the IC energy-loss rate from the same kernel on a per-electron grid of 400
  outgoing energies (``cascade.py:107-138``, ``180-188``);
* gamma-gamma pair injection (Aharonian, Atoyan & Nagapetyan 1983 kernel,
  ``kernels.py:130-187``, tensor form ``cascade.py:141-171``);
* the Breit-Wheeler cross section and the opacity per unit length
  (``opacity.py:44-104``; field-free tensor ``cascade.py:425-441``);
* the kinetic single-scatter cascade iteration giving (T_gamma, T_electron)
  (``cascade.py:331-407``; the batched form ``444-589`` freezes each z at
  convergence exactly like the per-z loop);
* the Bethe-Heitler e+- source shape from the Blumenthal (1970) W-function
  (Kelner & Aharonian 2008 eq. 62), nested five-interval Simpson rules,
  normalised per proton column so the pair carries the proton energy
  (``bethe_heitler.py:35-351``).

Everything here is field-free setup (tensors built once per grid) plus a
per-redshift contraction with n(eps, z); the per-step use of T and R_BH runs
on the device.
"""

from __future__ import annotations

import numpy as np
from scipy.integrate import cumulative_trapezoid, trapezoid

from .cosmology import H
from .units import PRINCE_UNITS, EnergyGrid

M_E = PRINCE_UNITS.m_electron
M_P = PRINCE_UNITS.m_proton
C_CM = PRINCE_UNITS.c
SIGMA_T = 8.0 / 3.0 * np.pi * PRINCE_UNITS.r_electron**2
IC_PREF = 0.75 * SIGMA_T * C_CM

#: EM grid 1e-3 .. 1e14 GeV at 8 bins per decade (136 bins)
EM_GRID = (-3, 14, 8)
#: target-photon grids (cascade.py:286, 363; opacity grid 377-379; BH 587)
EPS_CASCADE = np.logspace(-15, -9, 400)
EPS_OPACITY = np.logspace(-12, -7, 256)
MU_OPACITY = np.linspace(-1.0, 1.0, 128)
EPS_BH = np.logspace(-12.5, -7.0, 40)
N_E1 = 400
COOL_FLOOR_FRAC = 0.1
MAX_SCATTER = 4000
CASCADE_TOL = 1e-4
BH_E_P_MIN = 1e5


def em_grid():
    return EnergyGrid(*EM_GRID)


def trapz_weights(x):
    """w with sum(w * y) == trapezoid(y, x)."""
    x = np.asarray(x, dtype=np.float64)
    w = np.empty_like(x)
    w[1:-1] = 0.5 * (x[2:] - x[:-2])
    w[0] = 0.5 * (x[1] - x[0])
    w[-1] = 0.5 * (x[-1] - x[-2])
    return w


def photon_density(field, eps, z):
    """n(eps, z) with non-finite and non-positive values set to 0."""
    n = np.asarray(field.get_photon_density(eps, z), dtype=np.float64)
    return np.where(np.isfinite(n) & (n > 0.0), n, 0.0)


def bg70_f(q, Gam):
    """Blumenthal & Gould (1970) eq. 2.48 spectral function f(q, Gamma)."""
    gq = Gam * q
    return 2.0 * q * np.log(q) + (1.0 + 2.0 * q) * (1.0 - q) + 0.5 * gq * gq * (1.0 - q) / (1.0 + gq)


def sigma_gg(s):
    """Breit-Wheeler gamma gamma -> e+ e- cross section [cm^2] at squared CM
    energy s [GeV^2]; zero at and below threshold (opacity.py:44-75).

    With w = 4 m_e^2 / s = 1 - beta^2 the bracket is written with
    (1 + beta)^2 / w in place of (1 + beta)/(1 - beta), finite as w -> 0.
    """
    s = np.asarray(s, dtype=np.float64)
    out = np.zeros_like(s)
    above = s > 4.0 * M_E * M_E
    w = 4.0 * M_E * M_E / s[above]
    b2 = 1.0 - w
    beta = np.sqrt(b2)
    out[above] = (3.0 / 16.0) * SIGMA_T * w * (
        (3.0 - b2 * b2) * np.log((1.0 + beta) ** 2 / w) - 2.0 * beta * (2.0 - b2)
    )
    return out


def bh_dsigma(k, g, u):
    """Born differential cross section d2sigma/(dgamma_e dxi) of Bethe-Heitler
    pair production (Blumenthal 1970 eq. 10, Kelner & Aharonian 2008 eq. 62)
    in units of (3/(16 pi)) sigma_T alpha.

    k: photon energy in the proton rest frame, g: electron Lorentz factor, u:
    cosine of the electron angle (all in m_e units).  Zero outside
    2.0001 <= k <= 600; denominators are floored at 1e-30
    (bethe_heitler.py:35-72, 139-166).
    """
    tiny = 1e-30
    fl = lambda v: np.maximum(v, tiny)  # noqa: E731
    with np.errstate(all="ignore"):
        p = np.sqrt(np.maximum(g * g - 1.0, 0.0) + tiny)
        g1 = k - g
        p1 = np.sqrt(np.maximum(g1 * g1 - 1.0, 0.0) + tiny)
        D = g - u * p
        T2 = np.maximum(k * k + p * p - 2.0 * k * p * u, 0.0)
        T = np.sqrt(T2)
        pp = p * p
        D2 = D * D
        log_T = np.log((T + p1) / fl(T - p1))
        y1 = np.log(fl((g1 + p1) / fl(g1 - p1))) / fl(p1)
        Y = 2.0 * np.log(fl((g * g1 + p * p1 + 1.0) / k)) / fl(pp)
        a_front = p * p1 / (k * k * k)
        terms = (
            -4.0 * (1.0 - u * u) * (2.0 * g * g + 1.0) / fl(pp * D2 * D2)
            + (5.0 * g * g - 2.0 * g * g1 + 3.0) / fl(pp * D2)
            + (pp - k * k) / fl(T2 * D2)
            + 2.0 * g1 / fl(pp * D)
        )
        inner = (
            2.0 * g * (1.0 - u * u) * (3.0 * k + pp * g1) / fl(D2 * D2)
            + (2.0 * g * g * (g * g + g1 * g1) - 7.0 * g * g - 3.0 * g * g1 - g1 * g1 + 1.0) / fl(D2)
            + k * (g * g - g * g1 - 1.0) / fl(D)
        )
        terms = terms + Y / fl(p * p1) * inner
        terms = terms - (log_T / fl(p1 * T)) * (
            2.0 / fl(D2) - 3.0 * k / fl(D) - k * (pp - k * k) / fl(T2 * D)
        )
        terms = terms - 2.0 * y1 / fl(D)
        res = a_front * terms
    ok = (k >= 2.0001) & (k <= 600.0) & (T >= tiny)
    return np.where(ok, res, 0.0)


def simpson5(fn, lo, hi, nint=5):
    """Composite Simpson rule with ``nint`` intervals over per-element [lo, hi]
    (zero where hi <= lo); the accumulation order of bethe_heitler.py:75-99."""
    valid = hi > lo
    dx = np.where(valid, (hi - lo) / nint, 0.0)
    x = np.array(lo, dtype=np.float64, copy=True)
    acc = np.zeros_like(x)
    for _ in range(nint):
        f0 = dx * fn(x)
        fm = dx * fn(x + 0.5 * dx)
        f1 = dx * fn(x + dx)
        acc = acc + (f0 + f1) / 6.0 + 2.0 * fm / 3.0
        x = x + dx
    return np.where(valid, acc, 0.0)


def bh_kernel(E_e, E_p, eps):
    """Field-free Bethe-Heitler electron kernel K[i, j, k] for electrons at
    E_e[i] from protons at E_p[j] on target photons eps[k] (shape only)
    (bethe_heitler.py:102-123, 169-214)."""
    ge = (np.asarray(E_e) / M_E)[:, None, None]
    gp = (np.asarray(E_p) / M_P)[None, :, None]
    ke = (np.asarray(eps) / M_E)[None, None, :]
    GE, GP, KE = np.broadcast_arrays(ge, gp, ke)
    om_hi = 2.0 * GP * KE
    om_lo = (GP + GE) ** 2 / (2.0 * GP * GE)
    ok = (om_hi - om_lo) / np.maximum(om_lo, 1e-30) >= 1e-8

    def over_omega(om):
        lo = (GP * GP + GE * GE) / (2.0 * GP * GE) + 1e-5
        hi = om - 1.0 - 1e-5
        good = ((hi - lo) >= -1e-4 * lo) & ((hi - lo) / lo >= 1e-4)

        def over_eps_e(ee):
            with np.errstate(all="ignore"):
                pe = np.sqrt(np.maximum(ee * ee - 1.0, 0.0) + 1e-20)
                xi = (GP * ee - GE) / (GP * pe)
                val = om * bh_dsigma(om, ee, xi) / pe
            inside = (pe >= 1e-20) & (xi >= -1.0) & (xi <= 1.0) & (ee >= 1.0) & (ee <= om - 1.0)
            return np.where(inside, val, 0.0)

        res = simpson5(over_eps_e, np.where(good, lo, 1.0), np.where(good, hi, 0.0))
        return np.where(good, res, 0.0)

    y = simpson5(over_omega, np.where(ok, om_lo, 1.0), np.where(ok, om_hi, 0.0))
    return np.where(ok, y * GE / (2.0 * GP**3 * KE**2), 0.0)


class CascadeKernels:
    """Field-free cascade tensors for one EM grid and one nucleon grid.

    ``transfer(zs, field, dz)`` returns the stiffness-split transfer pair
    (T_gamma, T_electron) at each z; ``bh_shape(zs, field)`` the
    Bethe-Heitler pair shapes.  Both contract the cached tensors with
    n(eps, z).
    """

    def __init__(self, E_em, E_cr):
        E = np.asarray(E_em, dtype=np.float64)
        self.E = E
        self.E_cr = np.asarray(E_cr, dtype=np.float64)
        self.n = n = E.size
        self.dE = np.gradient(E)
        eps = EPS_CASCADE
        weps = trapz_weights(eps)
        # -- IC emission A[i, j, k]: photon E[i] from electron E[j] on eps[k]
        g = E / M_E
        Gam = 4.0 * eps[None, None, :] * g[None, :, None] / M_E
        Eo = E[:, None, None]
        Ee = E[None, :, None]
        with np.errstate(all="ignore"):
            q = Eo / (Gam * (Ee - Eo))
            ok = (q >= 1.0 / (4.0 * g[None, :, None] ** 2)) & (q <= 1.0) & (Eo < Ee) & (g[None, :, None] > 1.0)
            self.A_emis = np.where(ok, bg70_f(q, Gam) / eps[None, None, :], 0.0) * weps[None, None, :]
        self.pref_emis = np.where(g > 1.0, IC_PREF / g**2, 0.0)
        del Gam, q, ok
        # -- IC loss: per electron j, outgoing grid E1[j, m] up to 0.9999 E[j]
        lo = np.log10(eps.min())
        self.E1 = np.stack([np.logspace(lo, np.log10(0.9999 * E[j]), N_E1) for j in range(n)])
        self.w1 = np.stack([trapz_weights(self.E1[j]) for j in range(n)])
        Gam = 4.0 * eps[None, None, :] * g[:, None, None] / M_E
        Eo = self.E1[:, :, None]
        Ee = E[:, None, None]
        with np.errstate(all="ignore"):
            q = Eo / (Gam * (Ee - Eo))
            ok = (q >= 1.0 / (4.0 * g[:, None, None] ** 2)) & (q <= 1.0) & (Eo < Ee) & (g[:, None, None] > 1.0)
            self.B_loss = np.where(ok, bg70_f(q, Gam) / eps[None, None, :], 0.0) * weps[None, None, :]
        del Gam, q, ok
        # -- pair injection A_p[i, j, k]: lepton E[i] from photon E[j] on eps[k]
        ge = (E / M_E)[:, None, None]
        eg = (E / M_E)[None, :, None]
        w = (eps / M_E)[None, None, :]
        den = ge * (eg - ge)
        with np.errstate(all="ignore"):
            wmin = eg / (4.0 * den)
            F = (
                4.0 * eg**2 / den * np.log(4.0 * w * den / eg)
                - 8.0 * eg * w
                + 2.0 * (2.0 * eg * w - 1.0) * eg**2 / den
                - (1.0 - 1.0 / (eg * w)) * eg**4 / den**2
            )
            ok = (ge > 1.0) & (ge < eg - 1.0) & (den > 0.0) & (w > wmin) & (wmin / w < 1.0)
            self.A_pair = np.where(ok, F / (w * w) / M_E, 0.0) * weps[None, None, :]
        self.pref_pair = IC_PREF / (E / M_E) ** 3
        del F, ok, wmin, den
        # -- opacity per unit length: K_tau[i, k] with dtau/dl = K_tau @ n(eps_o)
        omm = 1.0 - MU_OPACITY
        s = 2.0 * E[:, None, None] * EPS_OPACITY[None, :, None] * omm[None, None, :]
        sig = sigma_gg(s)
        ang = 0.5 * trapezoid(omm[None, None, :] * sig, MU_OPACITY, axis=2)
        self.K_tau = ang * trapz_weights(EPS_OPACITY)[None, :]
        del s, sig
        # -- degraded-electron cloud-in-cell map (z independent)
        Eep = E[None, :] - E[:, None]  # [k, j] = E[j] - E[k]
        self._cic_valid = Eep > E[0]
        b = np.clip(np.searchsorted(E, Eep.ravel()) - 1, 0, n - 2).reshape(n, n)
        self._cic_b = b
        self._cic_f = np.clip((E[b + 1] - Eep) / (E[b + 1] - E[b]), 0.0, 1.0)
        # -- Bethe-Heitler kernel (built on first use: the slowest tensor)
        self._K_bh = None

    # ------------------------------------------------------------ per z
    def _energy_normalise(self, M):
        """Scale each column j so that trapezoid(E * M[:, j], E) == E[j]
        (cascade.py:608-616).  The column is divided by its energy before the
        multiplication by E[j], so a column whose energy is denormal (a field
        that is ~exp(-700) at threshold) stays finite instead of overflowing."""
        out_E = trapezoid(self.E[:, None] * M, self.E, axis=0)
        pos = out_E > 0.0
        safe = np.where(pos, out_E, 1.0)
        return np.where(pos[None, :], (M / safe[None, :]) * self.E[None, :], M)

    def matrices(self, n_eps):
        """(P, Sg, De, Mic, R) at one photon density n(EPS_CASCADE)."""
        n, E, dE = self.n, self.E, self.dE
        emis = self.pref_emis[None, :] * (self.A_emis @ n_eps)
        pair = self.pref_pair[None, :] * (self.A_pair @ n_eps)
        P = self._energy_normalise(pair)
        R = (emis * dE[:, None]).sum(axis=0)
        Sg = np.where(R[None, :] > 0.0, emis / np.where(R > 0.0, R, 1.0)[None, :], 0.0)
        # degraded electrons: weight W[k, j] lands at E[j] - E[k], split between
        # the two bracketing bins so number and energy are both kept
        W = Sg * dE[:, None]
        valid = (R[None, :] > 0.0) & (W > 0.0) & self._cic_valid
        De = np.zeros((n, n))
        kk, jj = np.nonzero(valid)
        b = self._cic_b[kk, jj]
        f = self._cic_f[kk, jj]
        wv = W[kk, jj]
        np.add.at(De, (b, jj), wv * f / dE[b])
        np.add.at(De, (b + 1, jj), wv * (1.0 - f) / dE[b + 1])
        # IC cooling: loss rate and the cooled-photon matrix
        dN = self.pref_emis[:, None] * (self.B_loss @ n_eps)
        power = (self.w1 * self.E1 * dN).sum(axis=1)
        rate = (self.w1 * dN).sum(axis=1)
        we = trapz_weights(EPS_CASCADE)
        den = (we * n_eps).sum()
        eps_mean = (we * EPS_CASCADE * n_eps).sum() / den if den > 0 else 0.0
        loss = power - rate * eps_mean
        gcool = np.where(loss[None, :] > 0.0, emis / np.where(loss > 0.0, loss, 1.0)[None, :], 0.0)
        cum = cumulative_trapezoid(gcool, E, axis=1, initial=0.0)
        Mraw = cum - np.diag(cum)[:, None]
        Mraw = np.where(np.arange(n)[None, :] >= np.arange(n)[:, None], Mraw, 0.0)
        Mic = self._energy_normalise(Mraw)
        return P, Sg, De, Mic, R

    def escape(self, z, field, dz):
        """Per-step survival exp(-(dtau/dl) c |dz| / ((1+z) H(z)))."""
        no = photon_density(field, EPS_OPACITY, z)
        dtau = self.K_tau @ no
        dl = C_CM * abs(dz) / ((1.0 + z) * H(z))
        return np.exp(-dtau * dl)

    def transfer_one(self, z, field, dz, max_scatter=MAX_SCATTER, tol=CASCADE_TOL):
        """Kinetic single-scatter cascade at one z: (T_gamma, T_electron, iterations)."""
        E, dE = self.E, self.dE
        P, Sg, De, Mic, R = self.matrices(photon_density(field, EPS_CASCADE, z))
        pesc = self.escape(z, field, dz)[:, None]
        egam = (E[:, None] * Sg * dE[:, None]).sum(axis=0)
        cold = (egam < COOL_FLOOR_FRAC * E) | (R <= 0.0)
        hot = ~cold
        Ein = E * dE
        G = np.eye(self.n)
        e = np.zeros_like(G)
        esc = np.zeros_like(G)
        tot0 = Ein @ G
        it = 0
        for it in range(1, max_scatter + 1):
            esc += pesc * G
            e = e + P @ ((G - pesc * G) * dE[:, None])
            e_cold = e * cold[:, None]
            e_hot = e * hot[:, None]
            G_cold = Mic @ (e_cold * dE[:, None])
            G_hot = Sg @ (e_hot * dE[:, None])
            e = De @ (e_hot * dE[:, None])
            esc += pesc * G_cold
            G = G_hot + G_cold - pesc * G_cold
            if np.max((Ein @ G + Ein @ e) / tot0) < tol:
                esc += pesc * G
                break
        Te = esc @ (Mic * dE[None, :])
        return esc, Te, it

    def transfer(self, zs, field, dz):
        """Stacked (T_gamma, T_electron) over zs: arrays [Nz, n, n]."""
        Tg = np.empty((len(zs), self.n, self.n))
        Te = np.empty_like(Tg)
        self.iterations = []
        for i, z in enumerate(zs):
            Tg[i], Te[i], it = self.transfer_one(float(z), field, dz)
            self.iterations.append(it)
        return Tg, Te

    # ------------------------------------------------------------ Bethe-Heitler
    @property
    def K_bh(self):
        if self._K_bh is None:
            self._K_bh = bh_kernel(self.E, self.E_cr, EPS_BH)
        return self._K_bh

    def bh_shape_one(self, z, field):
        """R_BH[i, j]: one-sign e+- spectrum at E[i] per proton at E_cr[j],
        normalised so the pair carries E_cr[j] (bethe_heitler.py:219-265)."""
        E, Ep = self.E, self.E_cr
        n_eps = photon_density(field, EPS_BH, z)
        R = np.zeros((E.size, Ep.size))
        use = (n_eps > 0.0) & (EPS_BH > 0.0)
        if not np.any(use):
            return R
        ln_eps = np.log(EPS_BH[use])
        col = trapezoid(self.K_bh[:, :, use] * (n_eps * EPS_BH)[use][None, None, :], ln_eps, axis=2)
        band = (E[:, None] <= Ep[None, :]) & (E[:, None] > 1e-6 * Ep[None, :]) & (Ep[None, :] >= BH_E_P_MIN)
        col = np.where(band, col, 0.0)
        carried = 2.0 * trapezoid(E[:, None] * col, E, axis=0)
        good = carried > 0.0
        R[:, good] = col[:, good] * (Ep[good] / carried[good])[None, :]
        return R

    def bh_shape(self, zs, field):
        return np.stack([self.bh_shape_one(float(z), field) for z in zs])


_KERNEL_CACHE = {}


def cascade_kernels(E_em, E_cr):
    """Process-wide cache of CascadeKernels by grid."""
    key = (E_em.size, float(E_em[0]), float(E_em[-1]), E_cr.size, float(E_cr[0]), float(E_cr[-1]))
    k = _KERNEL_CACHE.get(key)
    if k is None:
        k = CascadeKernels(E_em, E_cr)
        _KERNEL_CACHE[key] = k
    return k
