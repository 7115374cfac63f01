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


def test_two_field_columns_match_oracle_with_scaled_ebl():
    run = workloads.nuclear_run()
    K = 3
    srcs = workloads.batch_sources(run, K, seed=5)
    scales = np.array([0.6, 1.0, 1.37])
    g = MultiRHSPropagationSolverETD2(0.12, 0.0, run, K=K)
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


def test_adjoint_gradient_matches_central_differences():
    prob = AdjointProblem(K=8, z_max=0.25)
    theta = draw_params(8, seed=7)
    losses, grad, tim = prob.loss_and_grad(theta)
    print("losses", losses, "timings", tim)
    # central differences of the GPU forward with the same knots and loss bins
    ip, mk = prob.last_loss_plan
    worst = 0.0
    for p in range(8):
        # relative step 1e-6 of each physical parameter: R_max itself for its log10
        h = 1e-6 / np.log(10.0) * np.ones(8) if p == 1 else 1e-6 * np.maximum(np.abs(theta[:, p]), 1.0)
        tp, tm = theta.copy(), theta.copy()
        tp[:, p] += h
        tm[:, p] -= h
        _, _, sp = prob.forward(tp, with_adjoint=False)
        lp = prob.loss_of_state(sp, ip, mk)[0].detach().cpu().numpy()
        _, _, sm = prob.forward(tm, with_adjoint=False)
        lm = prob.loss_of_state(sm, ip, mk)[0].detach().cpu().numpy()
        fd = (lp - lm) / (2 * h)
        err = np.abs(grad[:, p] - fd) / np.maximum(np.abs(fd), 1e-9 * np.abs(losses) / h + 1e-300)
        print(f"param {p}: adj {grad[:3, p]} fd {fd[:3]} max rel {err.max():.2e}")
        assert np.all(np.isfinite(fd)) and np.all(np.isfinite(grad[:, p]))
        worst = max(worst, float(err.max()))
    assert worst <= 1e-5
