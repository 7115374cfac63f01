"""GPU: per-column photon-field scales (two-field fold) and the adjoint gradient."""

import numpy as np
import pytest

from oracle.etd2_cpu import OracleETD2
from prince_amd import workloads
from prince_amd.adjoint import AdjointProblem, draw_params
from prince_amd.model.photonfields import CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL
from prince_amd.solvers.propagation import MultiRHSPropagationSolverETD2

from .parity import rel_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kt", [None, 8])  # default (2 per wave) and 8 per wave, 5 of them idle
def test_two_field_columns_match_oracle_with_scaled_ebl(kt):
    run = workloads.nuclear_run()
    K = 3
    srcs = workloads.batch_sources(run, K, seed=5)
    scales = np.array([0.6, 1.0, 1.37])
    g = MultiRHSPropagationSolverETD2(0.12, 0.0, run, K=K)
    if kt is not None:
        g._tuned = {("kt", K): kt}
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.set_field_scales(1, scales)
    g.solve(1e-3)
    base = run.photon_field
    amps = base.model_list[1].amplitudes
    try:
        for k in range(K):
            run.photon_field = CombinedPhotonField(
                [CMBPhotonSpectrum(), SyntheticEBL(amplitudes=amps, scale=scales[k])]
            )
            o = OracleETD2(0.12, 0.0, run)
            o.add_source_class(srcs[k])
            o.solve(1e-3)
            dev, nb = rel_dev(g[k].state, o.state)
            print(f"scale {scales[k]}: {nb} bins, max rel dev {dev:.3e}")
            assert dev <= 1e-8
    finally:
        run.photon_field = base


def _fd(fun, theta, p, h):
    """Richardson-extrapolated central difference (error O(h^4)) along parameter p."""

    def cd(step):
        tp, tm = theta.copy(), theta.copy()
        tp[:, p] += step
        tm[:, p] -= step
        return (fun(tp) - fun(tm)) / (2 * step)

    return (4 * cd(h / 2) - cd(h)) / 3


def _steps(theta, p):
    # relative step 1e-6 of each physical parameter; R_max itself for its log10
    if p == 1:
        return 1e-6 / np.log(10.0) * np.ones(theta.shape[0])
    return 1e-6 * np.maximum(np.abs(theta[:, p]), 1.0)


def test_adjoint_gradient_matches_central_differences():
    prob = AdjointProblem(K=8, z_max=0.25)
    theta = draw_params(8, seed=7)
    losses, grad, tim = prob.loss_and_grad(theta)
    print("losses", losses, "timings", tim)
    ip, mk = prob.last_loss_plan  # same knots and loss bins for every evaluation

    def fun(t):
        _, _, st = prob.forward(t, with_adjoint=False)
        return prob.loss_of_state(st, ip, mk)[0].detach().cpu().numpy()

    worst = 0.0
    for p in range(8):
        fd = _fd(fun, theta, p, _steps(theta, p))
        err = np.abs(grad[:, p] - fd) / np.abs(fd)
        print(f"param {p}: adj {grad[:, p]} fd {fd} rel {err}")
        assert np.all(np.isfinite(fd)) and np.all(np.isfinite(grad[:, p]))
        worst = max(worst, float(err.max()))
    assert worst <= 1e-5


def test_adjoint_gradient_matches_cpu_oracle_differences():
    """Adjoint on the GPU vs central differences of the CPU oracle (one instance)."""
    import torch

    from prince_amd.adjoint import make_source

    prob = AdjointProblem(K=1, z_max=0.06, target_seed=99)
    theta = draw_params(1, seed=21)
    _, grad, _ = prob.loss_and_grad(theta)
    ip, mk = prob.last_loss_plan
    run = prob.run
    base = run.photon_field
    amps = base.model_list[1].amplitudes

    def fun(t):
        th = t[0]
        run.photon_field = CombinedPhotonField(
            [CMBPhotonSpectrum(), SyntheticEBL(amplitudes=amps, scale=th[7])]
        )
        try:
            o = OracleETD2(0.06, 0.0, run)
            o.add_source_class(make_source(run, th))
            o.solve(1e-3)
        finally:
            run.photon_field = base
        st = torch.from_numpy(o.state.reshape(run.spec_man.nspec, run.cr_grid.d, 1)).to(mk.device)
        return prob.loss_of_state(st, ip, mk)[0].detach().cpu().numpy()

    for p in range(8):
        fd = _fd(fun, theta, p, _steps(theta, p))
        rel = abs(grad[0, p] - fd[0]) / abs(fd[0])
        print(f"param {p}: adj {grad[0, p]:.10e} cpu-fd {fd[0]:.10e} rel {rel:.2e}")
        assert np.isfinite(fd[0]) and rel <= 1e-5
