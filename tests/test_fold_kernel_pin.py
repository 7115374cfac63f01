"""Redistributed (x-differential) channels of the rate kernel pinned by a
direct 2-D quadrature.

The kernel row of mother bin i, daughter bin k holds, per photon bin j,

    max(0, (m_p / E_i)(dE_i / dE_k) I_ijk) - [self channel, k = i] (m_p / E_i) N_ij

where I_ijk is the integral of the response density rho(x, y) over the cell
x in [E_k,lo, E_k,hi] / E_i, y in gamma_i [eps_j,lo, eps_j,hi], and N_ij the
non-elastic integral of the mother (reference interaction_rates.py:488-595;
the non-elastic part is subtracted after the clip), kept only for
x_cut <= E_k,lo / E_i <= 1.  The response tables hold the double trapezoid
antiderivative of f(x, y) = int_0^{2y} eps sigma(x, eps) d eps / (2 y^2) on
(xcenters, ygrid), read bilinearly; a bilinear antiderivative has a
piecewise-constant mixed derivative, which for a double trapezoid is the mean
of f over the four corners of each table cell.  Here I_ijk is computed as
sum over table cells of (overlap area) x (corner mean of f) -- from the cross
section itself, with rectangle overlaps instead of corner sampling.
"""

import numpy as np
import pytest

from prince_amd import workloads
from prince_amd.model import config
from prince_amd.model.units import PDG_PROTON, PRINCE_UNITS, make_nucleus_pdg


def _overlap(lo, hi, edges):
    """[len(lo), len(edges) - 1] lengths of [lo, hi] inside each table interval."""
    a = np.maximum(lo[:, None], edges[None, :-1])
    b = np.minimum(hi[:, None], edges[None, 1:])
    return np.maximum(b - a, 0.0)


def _response_density(cs, mo, da, ygr):
    """Corner mean of f(x, y) on every (xcenters, ygrid) table cell."""
    eg, sig = cs.incl_diff(mo, da)  # sig [nx, ne]
    g = eg * sig
    cum = np.concatenate([np.zeros((g.shape[0], 1)),
                          np.cumsum(0.5 * (g[:, 1:] + g[:, :-1]) * np.diff(eg), axis=1)], axis=1)
    f = cum[:, 1:] / (2.0 * ygr**2)  # y = eps_edge / 2 on egrid[1:]
    return 0.25 * (f[:-1, :-1] + f[1:, :-1] + f[:-1, 1:] + f[1:, 1:])


def _rows_of(ir, run, mo, da):
    d = run.cr_grid.d
    r0 = run.spec_man.pdgid2sref[da].lidx()
    c0 = run.spec_man.pdgid2sref[mo].lidx()
    sel = ((ir._batch_rows >= r0) & (ir._batch_rows < r0 + d)
           & (ir._batch_cols >= c0) & (ir._batch_cols < c0 + d))
    idx = np.flatnonzero(sel)
    return {(int(ir._batch_rows[t] - r0), int(ir._batch_cols[t] - c0)): int(t) for t in idx}


@pytest.mark.parametrize("mo", [PDG_PROTON, make_nucleus_pdg(4, 2), make_nucleus_pdg(56, 26)])
def test_redistributed_tiles_against_cell_quadrature(mo):
    run = workloads.nuclear_run()
    ir, cs = run.int_rates, run.cross_sections
    resp = cs.resp
    da = PDG_PROTON
    d = run.cr_grid.d
    ecr, bcr, wcr = run.cr_grid.grid, run.cr_grid.bins, run.cr_grid.widths
    bph = run.ph_grid.bins
    m_p = PRINCE_UNITS.m_proton
    xc, ygr = resp.xcenters, resp.response_ygrid
    rho = _response_density(cs, mo, da, ygr)
    rows = _rows_of(ir, run, mo, da)
    B = ir._batch_matrix
    x_min = config.x_cut_proton  # proton daughters
    nonel = None
    if mo == da:
        f_non = resp.nonel_f_stack[resp.nonel_keys.index(mo)]
        nonel = f_non
    kept = clipped = neg_diag = 0
    for i in range(d):
        gam = ecr[i] / m_p
        y_lo, y_hi = gam * bph[:-1], gam * bph[1:]
        Oy = _overlap(y_lo, y_hi, ygr)  # [dph, ny-1]
        x_lo, x_hi = bcr[:-1] / ecr[i], bcr[1:] / ecr[i]
        Ox = _overlap(x_lo, x_hi, xc)  # [d, nx-1]
        I = Ox @ rho @ Oy.T  # [k, j]
        for k in range(d):
            keep = x_min <= x_lo[k] <= 1.0
            if not keep:
                assert (k, i) not in rows
                continue
            want = np.maximum((1.0 / gam) * (wcr[i] / wcr[k]) * I[k], 0.0)
            clipped += int(np.sum((I[k] < 0)))
            if nonel is not None and k == i:
                # exact integral of the piecewise-linear non-elastic response
                nn = np.zeros(bph.size - 1)
                for j in range(nn.size):
                    a, b = max(y_lo[j], ygr[0]), min(y_hi[j], ygr[-1])
                    if b > a:
                        pts = np.concatenate([[a], ygr[(ygr > a) & (ygr < b)], [b]])
                        nn[j] = np.trapezoid(np.interp(pts, ygr, nonel), pts)
                want = want - nn / gam
                # subtracted after the clip: the self-channel diagonal goes negative
                neg_diag += int(np.any(want < 0))
            got = B[rows[(k, i)]]
            scale = max(np.abs(want).max(), np.abs(got).max(), 1e-300)
            assert np.allclose(got, want, rtol=1e-9, atol=1e-11 * scale), (i, k)
            kept += 1
    assert kept > 100
    assert len(rows) == kept
    if nonel is not None:
        assert neg_diag > 10
    print(f"mother {mo}: {kept} kept rows match the cell quadrature ({clipped} negative cells)")


@pytest.mark.parametrize("which", ["proton", "nuclear"])
def test_toeplitz_factors_expand_to_the_dense_kernel(which):
    """The factored kernel (per-channel tables + row descriptors) expands to
    the dense kernel bit for bit, in a quarter of its bytes or less."""
    run = workloads.proton_run() if which == "proton" else workloads.nuclear_run()
    ir = run.int_rates
    tz = ir.toeplitz
    assert tz.desc.shape == (ir._batch_matrix.shape[0], 5)
    full = tz.expand()
    assert np.array_equal(full.view(np.int64), ir._batch_matrix.view(np.int64))
    rows = np.array([0, 7, full.shape[0] - 1])
    assert np.array_equal(tz.expand(rows), ir._batch_matrix[rows])
    if which == "nuclear":
        assert tz.nbytes < 0.26 * ir._batch_matrix.nbytes
