"""Photo-nuclear rate kernel and continuous loss rates.

Rate kernel.  With photon density n(eps, z) sampled at the photon-grid
centres, the coupling matrix M of the nuclear states has fixed CSR pattern
and values

    M.data = B @ n(eps, z)                     (interaction_rates.py:966-1067)

where each row of the dense FP64 kernel ``B`` belongs to one (daughter state,
mother state) entry.  Because the cosmic-ray centres, cosmic-ray edges and
photon edges share one logarithmic step, the integral of a response function
over a photon bin, seen from mother energy bin i, only depends on i + j
(photon bin j) -- and for redistributed daughters additionally on the bin
offset k - i of the daughter energy.  So per channel one table of
antiderivative differences on a log grid gives every row by shifting a
window (the reference's construction, ``interaction_rates.py:251-646``):

* boost-conserving daughter (or a mother's own non-elastic loss): mother bin
  i contributes the row (m_p / E_i) (dR_incl - dR_nonel)[i + j];
* redistributed daughter bin k of mother bin i: the row
  (m_p / E_i)(dE_i / dE_k) ddR[i + j, k - i + d - 1], negative values set to
  zero, the mother's non-elastic row subtracted afterwards when k = i and
  the daughter is the mother; only pairs with x_cut <= E_k,low / E_i <= 1
  are kept (x_cut_proton for proton daughters).

The 1-D antiderivative of a response function is that of its piecewise-
linear interpolant (zero outside the tabulated range), sampled on the y
grid (``interaction_rates.py:25-47``); the 2-D double antiderivative is
sampled bilinearly with zero fill (``util.py:106-160``).  Rows are ordered as
the entries of the canonical CSR ``coupling_mat`` (``interaction_rates.py:
648-697``); ``_batch_matrix``, ``_batch_rows``, ``_batch_cols`` and
``coupling_mat`` are the reference's attribute names for them.

Loss rates: adiabatic kappa = H(z) cm2sec E (``interaction_rates.py:
1194-1235``) and Bethe-Heitler pair production (``1238-1386``).
"""

from __future__ import annotations

import os

import numpy as np
from scipy.integrate import trapezoid
from scipy.interpolate import InterpolatedUnivariateSpline, RegularGridInterpolator
from scipy.sparse import csr_matrix

from . import config
from .cosmology import H
from .units import PDG_PROTON, PRINCE_UNITS, is_nucleus


def response_integral_operator(ygr, y_grid):
    """Matrix Q with (Q @ f)[m] = integral of the piecewise-linear
    interpolant of f (knots ygr, zero outside them) from ygr[0] to y_grid[m].

    Column k is that antiderivative for the unit vector e_k (the operator the
    reference builds from k = 1 splines, interaction_rates.py:25-47).
    """
    n = len(ygr)
    Q = np.empty((len(y_grid), n))
    unit = np.zeros(n)
    for k in range(n):
        unit[k] = 1.0
        spline = InterpolatedUnivariateSpline(ygr, unit, k=1, ext="zeros")
        Q[:, k] = spline.antiderivative()(y_grid)
        unit[k] = 0.0
    return Q


class _ResponseDifferences:
    """Per-channel differences of the response antiderivatives on the log grids.

    ``nonel[mo]`` / ``incl[(mo, da)]``: length d + dph - 1 vectors, entry
    m = R(y[m + 1]) - R(y[m]).  ``redist[(mo, da)]``: the second difference
    of the double antiderivative, shape (d + dph - 1, 2 d - 1), over
    (y cells, x cells).
    """

    def __init__(self, resp, y_grid, x_grid):
        Q = response_integral_operator(resp.response_ygrid, y_grid)
        self.nonel = self._first_differences(resp.nonel_keys, resp.nonel_f_stack, Q)
        self.incl = self._first_differences(resp.incl_keys, resp.incl_f_stack, Q)
        self.redist = {}
        keys = list(resp.incl_diff_integral_keys)
        if keys:
            sampler = RegularGridInterpolator(
                (resp.xcenters, resp.response_ygrid), resp.incl_diff_integral_stack,
                method="linear", bounds_error=False, fill_value=0.0,
            )
            yy, xx = np.meshgrid(y_grid, x_grid, indexing="ij")
            pts = np.column_stack([xx.ravel(), yy.ravel()])
            R2 = sampler(pts).reshape(y_grid.size, x_grid.size, len(keys))
            # corner sums of every (y cell, x cell): the integral over the cell
            cell = (R2[1:, 1:, :] - R2[1:, :-1, :]) - R2[:-1, 1:, :] + R2[:-1, :-1, :]
            for c, key in enumerate(keys):
                self.redist[key] = np.ascontiguousarray(cell[:, :, c])

    @staticmethod
    def _first_differences(keys, fstack, Q):
        if not len(keys):
            return {}
        diffs = np.diff(fstack @ Q.T, axis=1)
        return {key: diffs[row] for row, key in enumerate(keys)}


class ToeplitzFactors:
    """The rate kernel in factored form: per channel one table of response
    antiderivative differences, per kernel row a descriptor.

    Row r of ``_batch_matrix`` (CSR order) with descriptor
    ``desc[r] = (kind, t_gain, t_loss, i, k)`` is, for photon bin j,

    * kind BC:     fmo[i] tab1d[t_gain][i + j] - fmo[i] tab1d[t_loss][i + j]
    * kind REDIST: max(fdiff[i][k] tab2d[t_gain][i + j][k - i + d - 1], 0)
                   - fmo[i] tab1d[t_loss][i + j]

    where a term with table index -1 is absent (a BC row without gain is
    0 - fmo tab1d[t_loss]; the REDIST loss term only on the diagonal k = i of
    a self channel), fmo[i] = m_p / E_i and fdiff[i][k] = fmo[i] dE_i / dE_k.
    The same products in the same order as the dense construction, so the
    expansion reproduces ``_batch_matrix`` bit for bit.  Storage: the 1-D
    tables (d + dph - 1 values) and the 2-D tables ((d + dph - 1) x (2d - 1)
    values, redistributed channels only) instead of rows x dph values.
    """

    BC, REDIST = 0, 1

    def __init__(self, d, dph, fmo, fdiff):
        self.d, self.dph = int(d), int(dph)
        self.fmo = np.ascontiguousarray(fmo, dtype=np.float64)
        self.fdiff = np.ascontiguousarray(fdiff, dtype=np.float64)
        self._t1, self._t2 = [], []
        self._idx1, self._idx2 = {}, {}
        self.desc = None

    def table_1d(self, key, vec):
        if vec is None:
            return -1
        if key not in self._idx1:
            self._idx1[key] = len(self._t1)
            self._t1.append(np.asarray(vec, dtype=np.float64))
        return self._idx1[key]

    def table_2d(self, key, mat):
        if key not in self._idx2:
            self._idx2[key] = len(self._t2)
            self._t2.append(np.asarray(mat, dtype=np.float64))
        return self._idx2[key]

    @staticmethod
    def rows(kind, t_gain, t_loss, i, k):
        i = np.asarray(i, dtype=np.int32)
        out = np.empty((i.size, 5), dtype=np.int32)
        out[:, 0] = kind
        out[:, 1] = t_gain
        out[:, 2] = t_loss
        out[:, 3] = i
        out[:, 4] = np.asarray(k, dtype=np.int32)
        return out

    def finish(self):
        L = self.d + self.dph - 1
        self.tab1d = np.ascontiguousarray(np.stack(self._t1) if self._t1 else np.zeros((1, L)))
        self.tab2d = np.ascontiguousarray(
            np.stack(self._t2) if self._t2 else np.zeros((1, L, 2 * self.d - 1)))
        del self._t1, self._t2

    @property
    def nbytes(self):
        return self.tab1d.nbytes + self.tab2d.nbytes + self.desc.nbytes + self.fdiff.nbytes

    def expand(self, rows=None):
        """Host expansion of the listed kernel rows (all by default) [nrows, dph]."""
        desc = self.desc if rows is None else self.desc[np.asarray(rows)]
        d, dph = self.d, self.dph
        j = np.arange(dph)[None, :]
        kind, tg, tl, i, k = (desc[:, c] for c in range(5))
        out = np.zeros((desc.shape[0], dph))
        w = i[:, None] + j
        bc = kind == self.BC
        fm = self.fmo[i][:, None]
        t1 = np.where(bc & (tg >= 0), tg, 0)[:, None]
        t2 = np.where(bc, 0, tg)[:, None]
        gain_bc = np.where((tg >= 0)[:, None], fm * self.tab1d[t1, w], 0.0)
        red = self.fdiff[i, k][:, None] * self.tab2d[t2, w, (k - i + d - 1)[:, None]]
        red = np.maximum(red, 0.0)
        base = np.where(bc[:, None], gain_bc, red)
        has_loss = (tl >= 0) & (bc | (i == k))
        loss = fm * self.tab1d[np.maximum(tl, 0)[:, None], w]
        out[:] = np.where(has_loss[:, None], base - loss, base)
        return out


class PhotoNuclearInteractionRate:
    """Dense FP64 rate kernel ``_batch_matrix`` + fixed-pattern ``coupling_mat``."""

    def __init__(self, prince_run, kernel_cache=None):
        """``kernel_cache``: path of an ``.npz`` kernel cache (``kernel_cache.py``);
        loaded when it exists (and matches the run), written after the build
        otherwise."""
        from . import kernel_cache as kc

        self.prince_run = prince_run
        self.cross_sections = prince_run.cross_sections
        self.spec_man = prince_run.spec_man
        self.e_photon = prince_run.ph_grid
        self.e_cosmicray = prince_run.cr_grid
        self._ratemat_zcache = None
        self._check_common_log_step()
        if kernel_cache is not None and os.path.exists(kernel_cache):
            data = kc.load(prince_run, kernel_cache)
            n = prince_run.dim_states
            self.coupling_mat = csr_matrix(
                (np.zeros(data["indices"].size), data["indices"], data["indptr"]), shape=(n, n))
            self._batch_rows = data["batch_rows"]
            self._batch_cols = data["batch_cols"]
            self._batch_matrix = np.ascontiguousarray(data["batch_matrix"])
            self._batch_vec = np.zeros(self._batch_rows.size)
            self.sortidx = None
            self.toeplitz = None
            self.loaded_from_cache = True
            return
        rows, cols, vals = self._assemble_channels()
        self._store_in_csr_order(rows, cols, vals)
        self.loaded_from_cache = False
        if kernel_cache is not None:
            kc.save(self, kernel_cache)

    @property
    def photon_field(self):
        return self.prince_run.photon_field

    def photon_vector(self, z, pfield=None):
        """Photon density at the photon-grid centres (interaction_rates.py:105-116)."""
        field = self.photon_field if pfield is None else pfield
        return field.get_photon_density(self.e_photon.grid, z)

    # ------------------------------------------------------------ construction
    def _check_common_log_step(self):
        """The shifted-window construction needs one log step on all three grids."""
        steps = [np.log10(g[1] / g[0]) for g in
                 (self.e_cosmicray.grid, self.e_cosmicray.bins, self.e_photon.bins)]
        if not (np.isclose(steps[0], steps[1]) and np.isclose(steps[0], steps[2])):
            raise RuntimeError(
                "rate kernel: cosmic-ray centres, cosmic-ray edges and photon edges must share "
                "one log step (got %.6f, %.6f, %.6f)" % tuple(steps)
            )

    def _log_grids(self):
        """y grid (d + dph points) and x grid (2 d points) of the antiderivative tables."""
        ecr, bcr, bph = self.e_cosmicray.grid, self.e_cosmicray.bins, self.e_photon.bins
        d, dph = self.e_cosmicray.d, self.e_photon.d
        step = np.log10(ecr[1] / ecr[0])
        y_grid = (ecr[0] * bph[0] / PRINCE_UNITS.m_proton) * 10 ** (np.arange(d + dph) * step)
        x_grid = (bcr[0] / ecr[0]) * 10 ** ((np.arange(2 * d) - (d - 1)) * step)
        return y_grid, x_grid

    def _channels(self):
        """(mother, daughter, kind) of every channel of the run; kind is
        "bc" (boost-conserving or the mother's own non-elastic loss) or
        "redist".  Mothers are nuclei; both ends must be transported species."""
        cs = self.cross_sections
        species = set(self.spec_man.known_species)
        bc = cs.known_bc_channels_set
        redist = cs.known_diff_channels_set
        nonel_mothers = set(cs.resp.nonel_keys)
        pairs = set(bc) | set(redist) | {(m, m) for m in nonel_mothers}
        out = []
        for mo, da in sorted(pairs):
            if mo not in species or da not in species or not is_nucleus(mo):
                continue
            own_loss = mo == da and mo in nonel_mothers
            if (mo, da) in bc or (own_loss and (mo, da) not in redist):
                out.append((mo, da, "bc"))
            else:
                out.append((mo, da, "redist"))
        return out

    def _assemble_channels(self):
        d, dph = self.e_cosmicray.d, self.e_photon.d
        ecr, bcr = self.e_cosmicray.grid, self.e_cosmicray.bins
        width = self.e_cosmicray.widths
        y_grid, x_grid = self._log_grids()
        tabs = _ResponseDifferences(self.cross_sections.resp, y_grid, x_grid)

        ibin = np.arange(d)
        window = ibin[:, None] + np.arange(dph)[None, :]  # [mother bin, photon bin]
        mother_scale = PRINCE_UNITS.m_proton / ecr  # per mother bin
        # redistributed rows: [mother bin i, daughter bin k]
        redist_scale = mother_scale[:, None] * (width[:, None] / width[None, :])
        x_col = (ibin[None, :] - ibin[:, None]) + (d - 1)
        x_low = bcr[:-1][None, :] / ecr[:, None]
        mo_bin, da_bin = np.meshgrid(ibin, ibin, indexing="ij")

        sref = self.spec_man.pdgid2sref
        fac = ToeplitzFactors(d, dph, mother_scale, redist_scale)
        rows, cols, vals, descs = [], [], [], []
        for mo, da, kind in self._channels():
            r0, c0 = sref[da].lidx(), sref[mo].lidx()
            loss = tabs.nonel.get(mo) if mo == da else None
            t_loss = fac.table_1d(("nonel", mo), loss)
            if kind == "bc":
                gain = tabs.incl.get((mo, da))
                if gain is None and loss is None:
                    raise ValueError(f"rate kernel: channel {mo} -> {da} has no cross section")
                tile = mother_scale[:, None] * gain[window] if gain is not None else 0.0
                if loss is not None:
                    tile = tile - mother_scale[:, None] * loss[window]
                vals.append(np.broadcast_to(tile, (d, dph)).copy())
                rows.append(r0 + ibin)
                cols.append(c0 + ibin)
                descs.append(fac.rows(ToeplitzFactors.BC, fac.table_1d(("incl", mo, da), gain),
                                      t_loss, ibin, ibin))
                continue
            cube = redist_scale[:, :, None] * tabs.redist[(mo, da)][window[:, None, :], x_col[:, :, None]]
            np.maximum(cube, 0.0, out=cube)
            if loss is not None:
                cube[ibin, ibin, :] -= mother_scale[:, None] * loss[window]
            x_min = config.x_cut_proton if da == PDG_PROTON else config.x_cut
            keep = (x_low >= x_min) & (x_low <= 1)
            if not keep.any():
                continue
            vals.append(cube[keep])
            rows.append(r0 + da_bin[keep])
            cols.append(c0 + mo_bin[keep])
            descs.append(fac.rows(ToeplitzFactors.REDIST, fac.table_2d((mo, da), tabs.redist[(mo, da)]),
                                  t_loss, mo_bin[keep], da_bin[keep]))
        self._toeplitz_unsorted = fac
        fac.desc = np.concatenate(descs, axis=0)
        return (np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64),
                np.concatenate(vals, axis=0))

    def _store_in_csr_order(self, rows, cols, vals):
        """Sort the kernel rows into CSR order of (row, col) and build the pattern."""
        n = self.prince_run.dim_states
        order = np.lexsort((cols, rows))
        rows, cols = rows[order], cols[order]
        if rows.size > 1 and np.any((np.diff(rows) == 0) & (np.diff(cols) == 0)):
            raise RuntimeError("rate kernel: two channels write the same matrix entry")
        indptr = np.searchsorted(rows, np.arange(n + 1), side="left")
        self.coupling_mat = csr_matrix((np.zeros(rows.size), cols, indptr), shape=(n, n))
        self.sortidx = order
        self._batch_rows = rows
        self._batch_cols = cols
        self._batch_matrix = np.ascontiguousarray(vals[order, :])
        self._batch_vec = np.zeros(rows.size)
        fac = self.__dict__.pop("_toeplitz_unsorted", None)
        if fac is not None:
            fac.desc = np.ascontiguousarray(fac.desc[order])
            fac.finish()
        self.toeplitz = fac

    # ------------------------------------------------------------ evaluation
    def get_hadr_jacobian(self, z, scale_fac=1.0, force_update=False, pfield=None):
        """Coupling matrix with values folded at redshift z (interaction_rates.py:1066-1120)."""
        if force_update or pfield is not None or self._ratemat_zcache != z:
            np.dot(self._batch_matrix, self.photon_vector(z, pfield=pfield), out=self._batch_vec)
            self._ratemat_zcache = None if pfield is not None else z
            if scale_fac == 1.0:
                self.coupling_mat.data[:] = self._batch_vec
            else:
                np.multiply(self._batch_vec, scale_fac, out=self.coupling_mat.data)
        return self.coupling_mat


def _state_energies(prince_run, energy):
    """(energy per state, slice bounds per species) on grid centres or bin edges."""
    sm, grid = prince_run.spec_man, prince_run.cr_grid
    if energy == "grid":
        return grid.grid, prince_run.dim_states, (lambda s: (s.lidx(), s.uidx()))
    if energy == "bins":
        return grid.bins, prince_run.dim_bins, (lambda s: (s.lbin(), s.ubin()))
    raise ValueError(f"energy must be 'grid' or 'bins', got {energy!r}")


class ContinuousAdiabaticLossRate:
    """kappa_adia(z) = H(z) cm2sec E on every species block."""

    def __init__(self, prince_run, energy="grid"):
        self.spec_man = prince_run.spec_man
        self.e_cosmicray = prince_run.cr_grid
        earr, dim, bounds = _state_energies(prince_run, energy)
        self.energy_vector = np.zeros(dim)
        for spec in self.spec_man.species_refs:
            lo, hi = bounds(spec)
            self.energy_vector[lo:hi] = earr

    def loss_vector(self, z, energy=None):
        rate = H(z) * PRINCE_UNITS.cm2sec
        return rate * (self.energy_vector if energy is None else energy)


#: coefficients of phi(xi) (interaction_rates.py:1346-1386): the rational fit
#: below xi = 25 in u = xi - 2, Blumenthal's cubic in ln xi and its 1/xi series
PHI_LOW = (0.8048, 0.1459, 1.137e-3, -3.879e-6)
PHI_HIGH = (2.667, -14.45, 50.95, -86.07)
PHI_HIGH_SERIES = (2.91, 78.35, 1837)


def blumenthal_phi(xi, low=PHI_LOW, high=PHI_HIGH, series=PHI_HIGH_SERIES):
    """phi(xi) of Bethe-Heitler energy loss (Blumenthal 1970), as the reference
    evaluates it: below xi = 25 (pi/12) u^4 / S(u) with S(u) = c1 u + c2 u^2 +
    c3 u^3 + c4 u^4, u = xi - 2 (the Chodorowski, Zdziarski & Sikora 1992 fit
    coefficients; their published denominator is 1 + S(u), see the tests),
    above it xi P(ln xi) / (1 - f1/xi - f2/xi^2 - f3/xi^3) with Blumenthal's
    ultra-relativistic cubic P."""
    xi = np.asarray(xi, dtype=np.float64)
    c1, c2, c3, c4 = low
    a3, a2, a1, a0 = high
    f1, f2, f3 = series
    out = np.zeros(xi.shape)
    below = xi < 25.0
    u = xi[below] - 2
    S = c1 * u**1 + c2 * u**2 + c3 * u**3 + c4 * u**4
    out[below] = np.pi / 12.0 * u**4 / S
    v = xi[~below]
    t = np.log(v)
    P = ((a3 * t + a2) * t + a1) * t + a0
    out[~below] = (v * P) / (1 - f1 * v**-1 - f2 * v**-2 - f3 * v**-3)
    return out


class ContinuousPairProductionLossRate:
    """Bethe-Heitler pair-production loss rate of nuclei (interaction_rates.py:1238-1386).

    loss_vector(z) = scale_vec * r(E, z) on every nucleus block with
    r = trapz(n(m_e xi / (2 gamma), z) phi(xi) / xi^2, xi) over ``xi_steps``
    log-spaced xi in [2 + 1e-8, 1e16] (gamma = E / m_p per nucleon) and
    scale_vec = alpha r_e^2 m_e^2 Z^2 / A.  ``xi``, ``phi_xi2``,
    ``photon_grid``, ``photon_vector`` and ``scale_vec`` carry the
    reference's names.
    """

    def __init__(self, prince_run, energy="grid", xi_steps=400):
        self.prince_run = prince_run
        self.spec_man = prince_run.spec_man
        self.e_cosmicray = prince_run.cr_grid
        self.xi = np.logspace(np.log10(2 + 1e-8), 16.0, xi_steps)
        self.phi_xi2 = blumenthal_phi(self.xi) / (self.xi**2)
        earr, dim, bounds = _state_energies(prince_run, energy)
        self._bounds = bounds
        self._loss_dim = dim
        units = PRINCE_UNITS.fine_structure * PRINCE_UNITS.r_electron**2 * PRINCE_UNITS.m_electron**2
        self.scale_vec = np.zeros(dim)
        for spec in self.spec_man.species_refs:
            if spec.is_nucleus:
                lo, hi = bounds(spec)
                self.scale_vec[lo:hi] = units * abs(spec.charge) ** 2 / float(spec.A)
        # target-photon energy of every (nucleon energy, xi) sample
        gamma = earr / PRINCE_UNITS.m_proton
        self.photon_grid = np.outer(1 / gamma, self.xi) * PRINCE_UNITS.m_electron / 2.0
        flat = self.photon_grid.reshape(-1)
        self._ascending = flat.argsort()
        self._ascending_energies = flat[self._ascending]

    @property
    def photon_field(self):
        return self.prince_run.photon_field

    def photon_vector(self, z, pfield=None):
        """n(photon_grid, z): the field is evaluated once on the ascending energies."""
        field = self.photon_field if pfield is None else pfield
        out = np.zeros_like(self.photon_grid)
        out.reshape(-1)[self._ascending] = field.get_photon_density(self._ascending_energies, z)
        return out

    def loss_vector(self, z, pfield=None):
        per_energy = trapezoid(self.photon_vector(z, pfield=pfield) * self.phi_xi2, self.xi, axis=1)
        full = np.zeros(self._loss_dim)
        for spec in self.spec_man.species_refs:
            if spec.is_nucleus:
                lo, hi = self._bounds(spec)
                full[lo:hi] = per_energy
        return self.scale_vec * full
