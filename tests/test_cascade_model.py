"""EM cascade operators pinned by physics that does not come from the code under
test: the Breit-Wheeler peak, an independent quadrature of the angle-averaged
opacity, the Thomson limit of inverse Compton, the Racah high-energy
Bethe-Heitler total cross section, energy conservation of the transfer and of
the per-step hook, and the normalisation contract of the pair shape."""

import numpy as np
import pytest
from scipy.integrate import quad, trapezoid

from oracle.etd2_cpu import OracleETD2
from prince_amd import workloads
from prince_amd.model import cascade as C
from prince_amd.model.photonfields import CMBPhotonSpectrum
from prince_amd.model.units import PRINCE_UNITS as U

CMB = CMBPhotonSpectrum()


@pytest.fixture(scope="module")
def kernels():
    run = workloads.proton_em_run()
    return C.cascade_kernels(run.em_grid.grid, run.cr_grid.grid)


def test_breit_wheeler_peak_and_threshold():
    sth = 4.0 * C.M_E**2
    assert np.all(C.sigma_gg(np.array([0.5 * sth, sth])) == 0.0)
    s = np.logspace(np.log10(sth * 1.0001), np.log10(sth * 100), 20001)
    sg = C.sigma_gg(s)
    # Gould & Schreder (1967): maximum 1.70e-25 cm^2 near s = 2 s_th
    assert abs(sg.max() / 1.70e-25 - 1.0) < 5e-3
    assert 1.9 < s[sg.argmax()] / sth < 2.05
    # high energy: sigma -> (3/8) sigma_T (s_th/s) (ln(4 s/s_th) - 1) ... within 1 %
    big = sth * 1e6
    approx = 0.375 * C.SIGMA_T * (sth / big) * (np.log(4.0 * big / sth) - 1.0)
    assert abs(C.sigma_gg(np.array([big]))[0] / approx - 1.0) < 1e-2


def test_opacity_kernel_matches_angle_quadrature(kernels):
    E = kernels.E
    w = C.trapz_weights(C.EPS_OPACITY)
    for i, k in ((110, 10), (120, 60), (130, 200)):
        ref, _ = quad(lambda mu: 0.5 * (1 - mu) * C.sigma_gg(np.array([2 * E[i] * C.EPS_OPACITY[k] * (1 - mu)]))[0],
                      -1.0, 1.0, limit=400, epsabs=0.0, epsrel=1e-10)
        got = kernels.K_tau[i, k] / w[k]
        # 128-point trapezoid in mu (the kernel's own rule) against adaptive quadrature
        assert ref > 0 and abs(got / ref - 1.0) < 5e-3, (i, k, got, ref)


def test_inverse_compton_thomson_limit(kernels):
    """-dE/dt = 4/3 sigma_T c gamma^2 U and the scattering rate sigma_T c N for
    gamma eps << m_e."""
    eps = C.EPS_CASCADE
    n = C.photon_density(CMB, eps, 0.0)
    U_rad = trapezoid(eps * n, eps)
    N_ph = trapezoid(n, eps)
    assert abs(U_rad * 1e9 / 0.2604 - 1.0) < 5e-3  # 0.26 eV cm^-3 at T = 2.725 K
    dN = kernels.pref_emis[:, None] * (kernels.B_loss @ n)
    power = (kernels.w1 * kernels.E1 * dN).sum(axis=1)
    rate = (kernels.w1 * dN).sum(axis=1)
    loss = power - rate * U_rad / N_ph
    g = kernels.E / C.M_E
    for j in (10, 20, 30):
        assert abs(loss[j] / (4.0 / 3.0 * C.SIGMA_T * U.c * g[j] ** 2 * U_rad) - 1.0) < 3e-3
        assert abs(rate[j] / (C.SIGMA_T * U.c * N_ph) - 1.0) < 1e-2
    # Klein-Nishina suppression sets in at higher gamma
    assert loss[80] < 0.01 * 4.0 / 3.0 * C.SIGMA_T * U.c * g[80] ** 2 * U_rad


def test_pair_kernel_rate_tracks_breit_wheeler(kernels):
    """The pair-injection kernel's total rate follows c dtau/dl (full CMB band)
    in its energy dependence; its absolute scale cancels in the energy-normalised
    pair matrix of the cascade."""
    n = C.photon_density(CMB, C.EPS_CASCADE, 0.0)
    pair = kernels.pref_pair[None, :] * (kernels.A_pair @ n)
    rate = (pair * kernels.dE[:, None]).sum(axis=0)
    eps = np.logspace(-15, -9, 1500)
    ne = C.photon_density(CMB, eps, 0.0)
    mu = np.linspace(-1.0, 1.0, 2001)
    ratios = []
    for j in (100, 110, 120, 130):
        s = 2 * kernels.E[j] * eps[:, None] * (1 - mu[None, :])
        ang = 0.5 * trapezoid((1 - mu)[None, :] * C.sigma_gg(s), mu, axis=1)
        ratios.append(rate[j] / (U.c * trapezoid(ne * ang, eps)))
    ratios = np.array(ratios)
    assert ratios.max() / ratios.min() < 1.12, ratios


@pytest.mark.parametrize("k,tol", [(100.0, 6e-3), (300.0, 1.5e-3), (590.0, 1e-3)])
def test_bethe_heitler_cross_section_racah_limit(k, tol):
    """Integrating the W-function over electron energy and angle gives the total
    pair cross section (28/9) alpha r_e^2 (ln 2k - 109/42) at k >> 1."""
    ge = np.unique(np.concatenate([1 + np.logspace(-8, np.log10(k / 2 - 1), 1200),
                                   k - 1 - np.logspace(np.log10(k / 2 - 1), -8, 1200)]))
    t = np.logspace(-14, np.log10(2.0), 2400)
    xi = np.unique(np.concatenate([1 - t, -1 + t]))
    xi = xi[(xi >= -1) & (xi <= 1)]
    G, X = np.meshgrid(ge, xi, indexing="ij")
    d = C.bh_dsigma(np.full_like(G, k), G, X)
    tot = 3.0 / (16.0 * np.pi) * C.SIGMA_T * U.fine_structure * trapezoid(trapezoid(d, xi, axis=1), ge)
    racah = 28.0 / 9.0 * U.fine_structure * U.r_electron**2 * (np.log(2 * k) - 109.0 / 42.0)
    assert abs(tot / racah - 1.0) < tol


def test_bh_shape_normalisation(kernels):
    R = kernels.bh_shape_one(0.5, CMB)
    E, Ep = kernels.E, kernels.E_cr
    assert np.all(R >= 0.0)
    carried = 2.0 * trapezoid(E[:, None] * R, E, axis=0)
    active = carried > 0
    assert active.sum() > 20
    assert np.allclose(carried[active], Ep[active], rtol=1e-12)
    assert not np.any(active & (Ep < C.BH_E_P_MIN))
    # leptons never above the proton energy
    assert np.all(R[E[:, None] > Ep[None, :]] == 0.0)


def test_transfer_conserves_energy_and_passes_soft_photons(kernels):
    Tg, Te, it = kernels.transfer_one(0.3, CMB, 1e-3)
    E, dE = kernels.E, kernels.dE
    Ein = E * dE
    # electrons below ~1 TeV radiate part of their IC photons under the grid's
    # lower edge (1 MeV), so T_electron conserves energy only above that
    for T, keep in ((Tg, E > 0), (Te, E > 1e3)):
        assert np.all(T >= 0.0)
        cons = (Ein @ T) / Ein
        assert np.all(np.abs(cons[keep] - 1.0) < 2e-3), (cons[keep].min(), cons[keep].max())
    assert 1 < it < C.MAX_SCATTER
    soft = E < 10.0  # far below the pair threshold on the CMB: photons pass unchanged
    assert np.allclose(np.diag(Tg)[soft], 1.0, atol=1e-9)
    assert np.allclose(Tg[np.ix_(soft, soft)] - np.diag(np.diag(Tg)[soft]), 0.0, atol=1e-12)


def test_oracle_hook_moves_the_pair_energy_into_photons():
    """One EM hook: photon energy out = photon energy in + pair energy injected
    by the active proton bins (T conserves energy to its convergence
    tolerance), and e+- end at zero."""
    run = workloads.proton_em_run()
    o = OracleETD2(0.02, 0.0, run)
    o.add_source_class(workloads.proton_source(run))
    o._em_dz = 1e-3
    o._M_off = None
    o._z_cached = None
    o._kappa_pair = None
    o._operator_at(0.02)
    rng = np.random.default_rng(1)
    state = np.abs(rng.standard_normal(run.dim_states)) * 1e3
    d_em = run.em_grid.d
    em = {p: np.zeros(d_em) for p in (-11, 11, 22)}
    em[22] = np.abs(rng.standard_normal(d_em)) * (run.em_grid.grid > 1e3) * 1e-6
    E, dE = run.em_grid.grid, np.gradient(run.em_grid.grid)
    e_before = np.sum(E * dE * em[22])
    Ep = run.cr_grid.grid
    dEp = np.gradient(Ep)
    loss = run.pair_loss_rates_grid.loss_vector(0.02)
    from prince_amd.model.cosmology import H

    dl = U.c * 1e-3 / (1.02 * H(0.02))
    # proton bins whose pair shape is active (the Born cap omega <= 600 empties
    # the top bins: their loss is not re-injected, as in the reference)
    active = o._em_bh.sum(axis=0) > 0
    assert 20 < active.sum() < Ep.size
    injected = sum(A * np.sum((loss[sl] * dl * dEp * state[sl])[active]) for sl, A in o.bh_species)
    o._em_hook(state, em)
    e_after = np.sum(E * dE * em[22])
    assert injected > 0
    assert abs((e_after - e_before) / injected - 1.0) < 3e-3
    assert not np.any(em[11]) and not np.any(em[-11])
