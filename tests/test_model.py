"""CPU checks of the host-side model: grids, FD operator, photon fields,
loss rates, the rate-kernel construction and sources, against closed forms
and independent quadrature."""

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import InterpolatedUnivariateSpline

from prince_amd import workloads
from prince_amd.model import config
from prince_amd.model.cosmology import H
from prince_amd.model.cr_sources import AugerFitSource, BrokenEvolutionSource, SimpleSource
from prince_amd.model.interaction_rates import (
    ContinuousPairProductionLossRate,
    blumenthal_phi,
    response_integral_operator,
)
from prince_amd.model.partial_diff import DifferentialOperator, single_block
from prince_amd.model.photonfields import (
    PLANCK_PREF,
    CMBPhotonSpectrum,
    FlatPhotonSpectrum,
    SyntheticEBL,
)
from prince_amd.model.units import (
    PRINCE_UNITS,
    EnergyGrid,
    get_AZN,
    is_nucleus,
    make_nucleus_pdg,
)
from prince_amd.solvers.result import UHECRPropagationResult


def test_grids_and_pdg():
    g = EnergyGrid(*config.cosmic_ray_grid)
    p = EnergyGrid(*config.photon_grid)
    assert g.d == 64 and p.d == 72
    assert np.isclose(np.log10(g.bins[1] / g.bins[0]), 1 / 8)
    assert get_AZN(make_nucleus_pdg(56, 26)) == (56, 26, 30)
    assert make_nucleus_pdg(1, 1) == 2212 and get_AZN(2112) == (1, 0, 1)
    assert is_nucleus(2212) and not is_nucleus(22)


def test_fd_operator_exact_on_polynomials_in_lnE():
    grid = EnergyGrid(6, 14, 8)
    D = single_block(grid)
    lnE = np.log(grid.grid)
    d = grid.d
    # D f = (1/E) d f / d lnE ; bulk 5-point stencil exact to degree 4, the
    # 4-point rows to degree 3 and the 3-point edge rows to degree 2.  The edge
    # rows 0 and d-1 use offsets (1,2,3) / (-3,-2,-1): they differentiate at the
    # neighbouring node (no diagonal entry), scaled by their own 1/E.
    c = lnE[d // 2]
    for p in range(0, 5):
        f = (lnE - c) ** p
        deriv = p * (lnE - c) ** max(p - 1, 0) if p else np.zeros(d)
        exact = deriv / grid.grid
        exact[0] = deriv[1] / grid.grid[0]
        exact[d - 1] = deriv[d - 2] / grid.grid[d - 1]
        got = D @ f
        rows = {4: range(3, d - 3), 3: list(range(1, 3)) + list(range(d - 3, d - 1))}
        rows[2] = [0, d - 1]
        for deg, rr in rows.items():
            if p <= deg:
                for r in rr:
                    assert abs(got[r] - exact[r]) <= 1e-9 * (np.abs(D[r]).sum() * np.abs(f).max() + 1e-300)
    # rows 0 and d-1 have no diagonal entry
    assert D[0, 0] == 0.0 and D[d - 1, d - 1] == 0.0
    op = DifferentialOperator(grid, 3).operator
    assert op.shape == (3 * d, 3 * d) and op.nnz == 3 * np.count_nonzero(D)


def test_cmb_closed_forms():
    cmb = CMBPhotonSpectrum()
    kT = config.E_CMB
    E = np.logspace(np.log10(kT) - 6, np.log10(kT) + 3, 20001)
    for z in (0.0, 1.0, 3.0):
        n = cmb.get_photon_density(E * (1 + z), z)
        # u = pref * pi^4/15 * (kT(1+z))^4
        u = np.trapezoid(E * (1 + z) * n, E * (1 + z))
        assert abs(u / (PLANCK_PREF * np.pi**4 / 15 * (kT * (1 + z)) ** 4) - 1) < 1e-4
    # zero where E/(kT (1+z)) >= 709
    assert cmb.get_photon_density(np.array([710 * kT]), 0.0)[0] == 0.0
    assert FlatPhotonSpectrum().get_photon_density(np.ones(3), 1.0)[0] == 4e12


def test_ebl_scaling_and_linearity():
    ebl = SyntheticEBL(seed=0)
    assert np.all((ebl.amplitudes >= 0.5) & (ebl.amplitudes <= 1.5))
    E = np.logspace(-14, -7, 50)
    n0 = ebl.get_photon_density(E, 0.0)
    assert np.allclose(ebl.get_photon_density(E, 1.0), 2**1.8 * n0, rtol=1e-14)
    ebl2 = SyntheticEBL(seed=0, scale=0.7)
    assert np.allclose(ebl2.get_photon_density(E, 0.5), 0.7 * ebl.get_photon_density(E, 0.5), rtol=1e-14)
    # stellar peak of E^2 n near 1 micron (1.24 eV), dust peak near 100 micron
    Ef = np.logspace(-13, -8, 4000)
    e2n = Ef**2 * SyntheticEBL(amplitudes=(1.0, 0.0)).density_z0(Ef)
    lam_um = 1.23984e-15 / Ef[np.argmax(e2n)] * 1e6 * 1e9 * 1e-9
    assert 0.5 < lam_um < 2.0


def test_hubble_and_dldz():
    z = 1.3
    assert H(z) == pytest.approx(config.H_0s * np.sqrt(0.27 * 2.3**3 + 0.73), rel=1e-15)
    run = workloads.proton_run()
    from prince_amd.solvers.propagation import UHECRPropagationSolver

    s = UHECRPropagationSolver(1.0, 0.0, run)
    assert s.dldz(z) == pytest.approx(-1.0 / (2.3 * H(z) * PRINCE_UNITS.cm2sec), rel=1e-15)
    assert s.dldz(0.0) < 0


def test_pair_loss_phi_positive_on_the_xi_grid():
    xs = np.logspace(np.log10(2 + 1e-8), 16, 400)
    assert np.all(blumenthal_phi(xs) > 0)


def test_response_integral_operator_is_antiderivative_of_linear_interpolant():
    rng = np.random.default_rng(0)
    ygr = np.sort(rng.uniform(1, 10, 15))
    f = rng.uniform(0, 2, 15)
    yq = np.concatenate([np.linspace(0.5, 11, 60), ygr])
    M = response_integral_operator(ygr, yq)
    got = M @ f

    def F(y):
        if y < ygr[0] or y > ygr[-1]:
            return 0.0  # zero extrapolation of the antiderivative spline
        return quad(lambda t: np.interp(t, ygr, f), ygr[0], y, points=list(ygr), limit=200)[0]

    exact = np.array([F(y) for y in yq])
    assert np.allclose(got, exact, rtol=1e-10, atol=1e-12)
    spl = InterpolatedUnivariateSpline(ygr, f, k=1, ext="zeros").antiderivative()
    assert np.allclose(got, spl(yq), rtol=1e-13, atol=1e-14)


def _channel_rows(ir, run, mo, da):
    d = run.cr_grid.d
    r0 = run.spec_man.pdgid2sref[da].lidx()
    c0 = run.spec_man.pdgid2sref[mo].lidx()
    out = {}
    for k, (r, c) in enumerate(zip(ir._batch_rows, ir._batch_cols)):
        if r0 <= r < r0 + d and c0 <= c < c0 + d:
            out[(r - r0, c - c0)] = k
    return out


def test_bc_and_nonel_tiles_against_quadrature():
    run = workloads.nuclear_run()
    ir = run.int_rates
    cs = run.cross_sections
    resp = cs.resp
    ygr = resp.response_ygrid
    mo = make_nucleus_pdg(28, 14)
    da = next(k[1] for k in cs.incl_idcs if k[0] == mo)
    m_p = PRINCE_UNITS.m_proton
    ecr = run.cr_grid.grid
    bph = run.ph_grid.bins
    f_incl = resp.incl_f_stack[resp.incl_keys.index((mo, da))]
    f_non = resp.nonel_f_stack[resp.nonel_keys.index(mo)]
    rows_bc = _channel_rows(ir, run, mo, da)
    rows_nd = _channel_rows(ir, run, mo, mo)
    for i in (5, 30, 60):
        gam = ecr[i] / m_p
        for j in (0, 20, 40, 71):
            lo, hi = gam * bph[j], gam * bph[j + 1]

            def integ(f):
                # exact integral of the piecewise-linear response: trapezoid over
                # the interval ends plus every knot inside it
                if hi < ygr[0] or lo > ygr[-1]:
                    return 0.0
                a, b = max(lo, ygr[0]), min(hi, ygr[-1])
                pts = np.concatenate([[a], ygr[(ygr > a) & (ygr < b)], [b]])
                return np.trapezoid(np.interp(pts, ygr, f), pts) / gam

            want_bc = integ(f_incl)
            got_bc = ir._batch_matrix[rows_bc[(i, i)], j]
            assert got_bc == pytest.approx(want_bc, rel=1e-7, abs=1e-40)
            want_nd = -integ(f_non)
            got_nd = ir._batch_matrix[rows_nd[(i, i)], j]
            assert got_nd == pytest.approx(want_nd, rel=1e-7, abs=1e-40)


def test_fold_is_linear_in_photon_field_and_csr_ordered():
    run = workloads.nuclear_run()
    ir = run.int_rates
    ph1 = ir.photon_vector(0.5)
    ph2 = CMBPhotonSpectrum().get_photon_density(run.ph_grid.grid, 2.0)
    fold = lambda ph: ir._batch_matrix @ ph  # noqa: E731
    assert np.allclose(fold(2 * ph1 + 3 * ph2), 2 * fold(ph1) + 3 * fold(ph2), rtol=1e-12,
                       atol=1e-300)
    cm = ir.coupling_mat
    rows = np.repeat(np.arange(cm.shape[0]), np.diff(cm.indptr))
    assert np.array_equal(rows, ir._batch_rows) and np.array_equal(cm.indices, ir._batch_cols)


def test_sources_closed_forms():
    run = workloads.nuclear_run()
    si = make_nucleus_pdg(28, 14)
    src = AugerFitSource(run, params={si: (1.0, 10**9.5, 2.0)}, norm=3.0, m=("simple", 2.0))
    E = run.cr_grid.grid
    ek = 28 * E
    emax = 10**9.5 * 14
    want = 2.0 * 28 * (ek / 1e9) ** -1.0 * np.where(ek < emax, 1.0, np.exp(1 - ek / emax))
    sl = run.spec_man.pdgid2sref[si].sl
    assert np.allclose(src.injection_grid[sl], want, rtol=1e-15)
    assert np.allclose(src.injection_rate(0.5), 3.0 * 1.5**2 * src.injection_grid, rtol=1e-15)
    b = BrokenEvolutionSource(run, params={si: (1.0, 10**9.5, 1.0)}, m=3.0, z_break=1.5)
    assert b.evolution(1.0) == 8.0 and b.evolution(2.9) == 2.5**3 and b.evolution(1.5) == 2.5**3
    p = SimpleSource(workloads.proton_run(), params={2212: (2.0, 1e11, 1.0)})
    Ep = workloads.proton_run().cr_grid.grid
    assert np.allclose(p.injection_grid, Ep**-2.0 * np.exp(-Ep / 1e11), rtol=1e-15)
    with pytest.raises(Exception):
        p.evolution(-0.1)


def test_injection_coefficients_are_the_pointwise_evolution_bit_for_bit():
    """The per-step coefficient table of a solve equals norm * evolution(z)
    evaluated point by point, bit for bit, for every evolution law (at the
    piecewise boundaries too) and for a subclass that overrides evolution."""
    run = workloads.nuclear_run()
    si = make_nucleus_pdg(28, 14)
    zs = np.concatenate([np.linspace(0.0, 5.0, 1001), [0.97, 1.0, 4.48, 1.5]])
    laws = ["flat", 2.3, ("SFR", 1.2), ("TDE", -1.0), ("simple", 2.7), ("simple_flat", 3.1),
            ("simple_SFR", 2.2)]
    srcs = [AugerFitSource(run, params={si: (1.0, 1e9, 1.0)}, norm=1.7, m=m) for m in laws]
    srcs.append(BrokenEvolutionSource(run, params={si: (1.0, 1e9, 1.0)}, norm=0.3, m=3.0))
    for src in srcs:
        want = np.array([src.norm * src.evolution(z) for z in zs])
        got = src.injection_coefficients(zs)
        assert got.dtype == np.float64 and np.array_equal(got, want), src.source_evo_m
    with pytest.raises(ValueError):
        srcs[1].injection_coefficients([0.1, -0.1])


def test_step_rates_are_the_pointwise_dldz_and_hubble_bit_for_bit():
    """The per-step tables (dl/dz and H cm2sec) a solve builds equal the scalar
    dldz and H evaluated point by point, bit for bit."""
    from prince_amd.model.cosmology import hubble_steps
    from prince_amd.solvers.propagation import UHECRPropagationSolver, z_grid_for

    solver = UHECRPropagationSolver.__new__(UHECRPropagationSolver)
    for zmax in (0.013, 0.73, 3.0, 4.97):
        z0 = z_grid_for(zmax, 0.0, 1e-3)[:-1]
        assert np.array_equal(np.array(hubble_steps(z0)), np.array([H(z) for z in z0]))
        dl, hub = solver.step_rates(z0)
        assert np.array_equal(dl, np.array([solver.dldz(z) for z in z0]))
        assert np.array_equal(hub, np.array([H(z) * PRINCE_UNITS.cm2sec for z in z0]))


def test_result_accessors_pure_species():
    run = workloads.nuclear_run()
    sm = run.spec_man
    fe = make_nucleus_pdg(56, 26)
    state = np.zeros(run.dim_states)
    state[sm.pdgid2sref[fe].sl] = run.cr_grid.grid ** -2.0
    res = UHECRPropagationResult(state, run.cr_grid.grid, sm)
    e, s = res.get_solution_scale(fe, epow=0)
    assert np.allclose(e, 56 * run.cr_grid.grid) and np.allclose(s, state[sm.pdgid2sref[fe].sl] / 56)
    egrid, mean, var = res.get_lnA([fe])
    ok = np.isfinite(mean)
    assert ok.any() and np.allclose(mean[ok], np.log(56)) and np.allclose(var[ok], 0, atol=1e-12)
    d = res.to_dict()
    back = UHECRPropagationResult.from_dict(d)
    assert np.array_equal(back.state, state)
    two = res + res * 2.0
    assert np.allclose(two.state, 3 * state)
