"""Photo-nuclear rate kernel and continuous loss rates.

``PhotoNuclearInteractionRate`` builds the dense FP64 fold kernel ``B``
(``batch_dim x dph``) whose product with the photon density gives the values
of the coupling matrix M on a fixed CSR pattern:

    M.data = B @ n_gamma(eps_ph, z)            (interaction_rates.py:966-1067)

Construction follows the reference's log-grid / Toeplitz recipe
(``interaction_rates.py:251-646``):

* boost-conserving channel tile  B[i_mo, j] = (m_p/E_mo) dR[i_mo + j]
  minus the non-elastic tile on the (mo, mo) diagonal;
* redistributed channel tile
  B[i_mo, i_da, j] = (m_p/E_mo)(dE_mo/dE_da) ddR[i_mo + j, i_da - i_mo + d - 1],
  clipped at 0 before the non-elastic subtraction, kept where
  x_cut <= bcr[i_da]/ecr[i_mo] <= 1 (x_cut_proton for proton daughters);
* rows are lexsorted by (row, col) so that row k of ``B`` feeds entry k of
  the canonical CSR ``coupling_mat.data`` (``interaction_rates.py:648-697``).

The 1-D antiderivatives use the same linear operator as the reference
(a k=1 zero-extrapolating spline's antiderivative sampled on the y grid,
``interaction_rates.py:25-47``); the 2-D ones use bilinear interpolation with
zero fill (``util.py:106-160``).  Both are evaluated with scipy, the library
the reference itself calls for them.

Loss rates: adiabatic kappa = H(z) cm2sec E (``1194-1235``) and the
Bethe-Heitler pair-production rate (Blumenthal phi(xi) fit, 400 log-spaced
xi, ``1238-1386``).
"""

from __future__ import annotations

import itertools

import numpy as np
from scipy.integrate import trapezoid
from scipy.interpolate import InterpolatedUnivariateSpline, RegularGridInterpolator
from scipy.sparse import csr_matrix

from . import config
from .cosmology import H
from .units import PDG_PROTON, PRINCE_UNITS, is_nucleus


def response_integral_operator(ygr, y_grid):
    """Matrix M with M @ f = antiderivative of the k=1 spline of f, on y_grid.

    Zero outside [ygr[0], ygr[-1]] (ext='zeros'), as the reference's
    ``_response_integral_operator`` (interaction_rates.py:25-47).
    """
    n = len(ygr)
    M = np.empty((len(y_grid), n))
    e = np.zeros(n)
    for k in range(n):
        e[k] = 1.0
        M[:, k] = InterpolatedUnivariateSpline(ygr, e, k=1, ext="zeros").antiderivative()(y_grid)
        e[k] = 0.0
    return M


class PhotoNuclearInteractionRate:
    """Dense FP64 fold kernel + fixed-pattern coupling matrix."""

    def __init__(self, prince_run):
        self.prince_run = prince_run
        self.cross_sections = prince_run.cross_sections
        self.spec_man = prince_run.spec_man
        self.e_photon = prince_run.ph_grid
        self.e_cosmicray = prince_run.cr_grid
        self._ratemat_zcache = None
        self._init_matrices()
        self._init_coupling_mat()

    @property
    def photon_field(self):
        return self.prince_run.photon_field

    def photon_vector(self, z, pfield=None):
        """Photon density at the photon-grid centres (interaction_rates.py:105-116)."""
        pf = pfield if pfield is not None else self.photon_field
        return pf.get_photon_density(self.e_photon.grid, z)

    # ------------------------------------------------------------------
    def _assert_log_grids_compatible(self):
        ecr, bcr, bph = self.e_cosmicray.grid, self.e_cosmicray.bins, self.e_photon.bins
        s_ecr = np.log10(ecr[1] / ecr[0])
        s_bcr = np.log10(bcr[1] / bcr[0])
        s_bph = np.log10(bph[1] / bph[0])
        if not (np.isclose(s_ecr, s_bcr) and np.isclose(s_ecr, s_bph)):
            raise RuntimeError(
                "Kernel construction requires the cosmic-ray and photon grids to share "
                f"the same log-step (got {s_ecr:.6f}, {s_bcr:.6f}, {s_bph:.6f})."
            )

    def _init_matrices(self):
        self._assert_log_grids_compatible()
        sp_id_ref = self.spec_man.pdgid2sref
        resp = self.cross_sections.resp
        m_pr = PRINCE_UNITS.m_proton
        ecr = self.e_cosmicray.grid
        bcr = self.e_cosmicray.bins
        bph = self.e_photon.bins
        dcr = self.e_cosmicray.d
        dph = self.e_photon.d
        delta_ec = self.e_cosmicray.widths

        log_step = np.log10(ecr[1] / ecr[0])
        y0 = ecr[0] * bph[0] / m_pr
        y_grid = y0 * 10 ** (np.arange(dcr + dph) * log_step)
        x0 = bcr[0] / ecr[0]
        x_grid = x0 * 10 ** ((np.arange(2 * dcr) - (dcr - 1)) * log_step)
        self._y_grid, self._x_grid = y_grid, x_grid

        factor_mo = m_pr / ecr
        factor_diff = factor_mo[:, None] * (delta_ec[:, None] / delta_ec[None, :])
        i_idx = np.arange(dcr)
        j_idx = np.arange(dph)
        A_2d = i_idx[:, None] + j_idx[None, :]
        B_idx = (i_idx[None, :] - i_idx[:, None]) + (dcr - 1)

        Mop = response_integral_operator(resp.response_ygrid, y_grid)
        nonel_dR, incl_dR = {}, {}
        if resp.nonel_f_stack.size:
            dRn = np.diff(resp.nonel_f_stack @ Mop.T, axis=1)
            nonel_dR = {mo: dRn[i] for i, mo in enumerate(resp.nonel_keys)}
        if resp.incl_f_stack.size:
            dRi = np.diff(resp.incl_f_stack @ Mop.T, axis=1)
            incl_dR = {key: dRi[i] for i, key in enumerate(resp.incl_keys)}

        diff_ddR = {}
        if resp.incl_diff_integral_keys:
            Y2d, X2d = np.meshgrid(y_grid, x_grid, indexing="ij")
            rgi = RegularGridInterpolator(
                (resp.xcenters, resp.response_ygrid),
                resp.incl_diff_integral_stack,
                method="linear",
                bounds_error=False,
                fill_value=0.0,
            )
            R_stack = rgi(np.column_stack([X2d.ravel(), Y2d.ravel()])).reshape(
                *Y2d.shape, len(resp.incl_diff_integral_keys)
            )
            ddR_stack = (
                R_stack[1:, 1:, :] - R_stack[1:, :-1, :] - R_stack[:-1, 1:, :] + R_stack[:-1, :-1, :]
            )
            for i, key in enumerate(resp.incl_diff_integral_keys):
                diff_ddR[key] = np.ascontiguousarray(ddR_stack[..., i])

        xl_2d = bcr[:-1][None, :] / ecr[:, None]
        diag = np.arange(dcr)
        imo_grid, ida_grid = np.meshgrid(i_idx, i_idx, indexing="ij")
        bc_set = self.cross_sections.known_bc_channels_set
        diff_set = self.cross_sections.known_diff_channels_set

        tiles, rows_l, cols_l = [], [], []
        known_species_rev = self.spec_man.known_species[::-1]
        for moid, daid in itertools.product(known_species_rev, known_species_rev):
            if not is_nucleus(moid):
                continue
            has_nonel = (moid == daid) and (moid in nonel_dR)
            in_bc = (moid, daid) in bc_set
            in_diff = (moid, daid) in diff_set
            if in_bc or (has_nonel and not in_diff):
                has_incl = (moid, daid) in incl_dR
                if not (has_nonel or has_incl):
                    raise Exception("Channel without interactions:", (moid, daid))
                tile = np.zeros((dcr, dph))
                if has_incl:
                    tile += factor_mo[:, None] * incl_dR[(moid, daid)][A_2d]
                if has_nonel:
                    tile -= factor_mo[:, None] * nonel_dR[moid][A_2d]
                tiles.append(tile)
                rows_l.append(sp_id_ref[daid].lidx() + i_idx)
                cols_l.append(sp_id_ref[moid].lidx() + i_idx)
            elif in_diff:
                ddR = diff_ddR[(moid, daid)]
                res = factor_diff[:, :, None] * ddR[A_2d[:, None, :], B_idx[:, :, None]]
                np.clip(res, 0.0, None, out=res)
                if has_nonel:
                    res[diag, diag, :] -= factor_mo[:, None] * nonel_dR[moid][A_2d]
                cut_low = config.x_cut_proton if daid == PDG_PROTON else config.x_cut
                cuts2d = np.logical_and(xl_2d >= cut_low, xl_2d <= 1)
                kept = res[cuts2d]
                if kept.shape[0] == 0:
                    continue
                tiles.append(kept)
                rows_l.append(sp_id_ref[daid].lidx() + ida_grid[cuts2d])
                cols_l.append(sp_id_ref[moid].lidx() + imo_grid[cuts2d])

        self._batch_matrix = np.concatenate(tiles, axis=0)
        self._batch_rows = np.concatenate(rows_l).astype(np.int64)
        self._batch_cols = np.concatenate(cols_l).astype(np.int64)

    def _init_coupling_mat(self):
        n = self.prince_run.dim_states
        nrows = self._batch_matrix.shape[0]
        self.coupling_mat = csr_matrix(
            (np.zeros(nrows), (self._batch_rows, self._batch_cols)), shape=(n, n), copy=True
        )
        if self.coupling_mat.nnz != nrows:
            raise RuntimeError("duplicate (row, col) entries in the rate kernel")
        self.sortidx = np.lexsort((self._batch_cols, self._batch_rows))
        self._batch_rows = self._batch_rows[self.sortidx]
        self._batch_cols = self._batch_cols[self.sortidx]
        self._batch_matrix = np.ascontiguousarray(self._batch_matrix[self.sortidx, :])
        self._batch_vec = np.zeros(nrows)

    # ------------------------------------------------------------------
    @property
    def batch_matrix(self):
        """Dense FP64 fold kernel, rows in canonical CSR order of ``coupling_mat``."""
        return self._batch_matrix

    def fold(self, photon):
        """M.data values for a given photon vector (one dense GEMV)."""
        return self._batch_matrix @ np.asarray(photon, dtype=np.float64)

    def get_hadr_jacobian(self, z, scale_fac=1.0, force_update=False, pfield=None):
        """Coupling matrix with values folded at redshift z (1066-1120)."""
        if force_update or pfield is not None or self._ratemat_zcache != z:
            np.dot(self._batch_matrix, self.photon_vector(z, pfield=pfield), out=self._batch_vec)
            self._ratemat_zcache = None if pfield is not None else z
            if scale_fac == 1.0:
                np.copyto(self.coupling_mat.data, self._batch_vec)
            else:
                np.multiply(self._batch_vec, scale_fac, out=self.coupling_mat.data)
        return self.coupling_mat


class ContinuousAdiabaticLossRate:
    """kappa_adia(z) = H(z) cm2sec E on every species block."""

    def __init__(self, prince_run, energy="grid"):
        self.spec_man = prince_run.spec_man
        self.e_cosmicray = prince_run.cr_grid
        if energy == "grid":
            earr, dim = self.e_cosmicray.grid, prince_run.dim_states
            lo, hi = (lambda s: s.lidx()), (lambda s: s.uidx())
        else:
            earr, dim = self.e_cosmicray.bins, prince_run.dim_bins
            lo, hi = (lambda s: s.lbin()), (lambda s: s.ubin())
        self.energy_vector = np.zeros(dim)
        for spec in self.spec_man.species_refs:
            self.energy_vector[lo(spec) : hi(spec)] = earr

    def loss_vector(self, z, energy=None):
        if energy is None:
            return H(z) * PRINCE_UNITS.cm2sec * self.energy_vector
        return H(z) * PRINCE_UNITS.cm2sec * energy


class ContinuousPairProductionLossRate:
    """Bethe-Heitler pair-production loss rate (interaction_rates.py:1238-1386).

    kappa_pair(z) = scale_vec * rate(E, z), with
    rate = trapz(n(m_e xi/(2 gamma), z) phi(xi)/xi^2, xi) and
    scale_vec = alpha r_e^2 m_e^2 Z^2/A on nucleus blocks.
    """

    def __init__(self, prince_run, energy="grid", xi_steps=400):
        self.prince_run = prince_run
        self.spec_man = prince_run.spec_man
        self.e_cosmicray = prince_run.cr_grid
        self.xi = np.logspace(np.log10(2 + 1e-8), 16.0, xi_steps)
        self.phi_xi2 = self._phi(self.xi) / (self.xi**2)
        if energy == "grid":
            earr, dim = self.e_cosmicray.grid, prince_run.dim_states
            self._lo, self._hi = (lambda s: s.lidx()), (lambda s: s.uidx())
        else:
            earr, dim = self.e_cosmicray.bins, prince_run.dim_bins
            self._lo, self._hi = (lambda s: s.lbin()), (lambda s: s.ubin())
        self._loss_dim = dim
        units = PRINCE_UNITS.fine_structure * PRINCE_UNITS.r_electron**2 * PRINCE_UNITS.m_electron**2
        self.scale_vec = np.zeros(dim)
        self.species_scale = np.zeros(self.spec_man.nspec)
        for spec in self.spec_man.species_refs:
            if not spec.is_nucleus:
                continue
            val = units * abs(spec.charge) ** 2 / float(spec.A)
            self.scale_vec[self._lo(spec) : self._hi(spec)] = val * np.ones_like(earr)
            self.species_scale[spec.princeidx] = val
        gamma = earr / PRINCE_UNITS.m_proton
        self.photon_grid = np.outer(1 / gamma, self.xi) * PRINCE_UNITS.m_electron / 2.0
        self.pg_desort = self.photon_grid.reshape(-1).argsort()
        self.pg_sorted = self.photon_grid.reshape(-1)[self.pg_desort]

    @property
    def photon_field(self):
        return self.prince_run.photon_field

    def photon_vector(self, z, pfield=None):
        pf = pfield if pfield is not None else self.photon_field
        pv = np.zeros_like(self.photon_grid)
        pv.reshape(-1)[self.pg_desort] = pf.get_photon_density(self.pg_sorted, z)
        return pv

    def rate_single(self, z, pfield=None):
        """Per-nucleon-energy rate shared by every nucleus block (length d)."""
        return trapezoid(self.photon_vector(z, pfield=pfield) * self.phi_xi2, self.xi, axis=1)

    def loss_vector(self, z, pfield=None):
        rs = self.rate_single(z, pfield=pfield)
        rate_full = np.zeros(self._loss_dim)
        for spec in self.spec_man.species_refs:
            if spec.is_nucleus:
                rate_full[self._lo(spec) : self._hi(spec)] = rs
        return self.scale_vec * rate_full

    @staticmethod
    def _phi(xi):
        """Blumenthal (1970) phi(xi) with the Chodorowski et al. fit below 25."""
        bltal_ultrarel = np.poly1d([2.667, -14.45, 50.95, -86.07])
        c1, c2, c3, c4 = 0.8048, 0.1459, 1.137e-3, -3.879e-6
        f1, f2, f3 = 2.91, 78.35, 1837
        res = np.zeros(xi.shape)
        le = np.where(xi < 25.0)
        he = np.where(xi >= 25.0)
        xl = xi[le] - 2
        res[le] = np.pi / 12.0 * xl**4 / (c1 * xl**1 + c2 * xl**2 + c3 * xl**3 + c4 * xl**4)
        xh = xi[he]
        res[he] = (xh * bltal_ultrarel(np.log(xh))) / (1 - f1 * xh**-1 - f2 * xh**-2 - f3 * xh**-3)
        return res
