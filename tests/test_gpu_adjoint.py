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
    """Adjoint on the GPU vs central differences of the CPU oracle (one
    instance, z 0.13 -> 0: 130 steps, so the sweep recomputes three
    checkpoint chunks of 64, 64 and 2 steps)."""
    import torch

    from prince_amd.adjoint import make_source

    prob = AdjointProblem(K=1, z_max=0.13, target_seed=99)
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
            o = OracleETD2(0.13, 0.0, run)
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


def _linearity_residuals(solver, plan, ab, lam_final, seed, log_coef=False):
    """<lam_final, x(dg)> against <gsrc, dg> per column for a random injection
    grid dg: the final state is linear in the injection grids (x0 = 0, the
    operator does not depend on them), so the two agree to rounding.  With
    ``log_coef`` the per-step source coefficients carry the extra ln(1+z) of
    the evolution-index derivative and the accumulator is gsrc1."""
    import torch

    b = plan["bufs"]
    grid0, coef0 = b["src_grid"].clone(), b["src_coef"].clone()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    dg = torch.randn(tuple(grid0.shape), generator=gen, dtype=torch.float64).to(grid0.device)
    acc = ab["gsrc1" if log_coef else "gsrc0"].clone()
    try:
        b["src_grid"].copy_(dg)
        if log_coef:
            lnz1 = torch.from_numpy(np.log1p(plan["z_grid"][:-1])).to(coef0.device)
            b["src_coef"].copy_(coef0 * lnz1[:, None, None])
        x = solver.execute(plan)  # [nspec, K, d]
        torch.cuda.synchronize()
        lhs = (lam_final * x).sum(dim=(0, 2))
        scale = (lam_final * x).abs().sum(dim=(0, 2))
        rhs = (acc * dg).sum(dim=(0, 1, 3))
    finally:
        b["src_grid"].copy_(grid0)
        b["src_coef"].copy_(coef0)
    lhs, rhs, scale = (t.cpu().numpy() for t in (lhs, rhs, scale))
    return np.abs(lhs - rhs) / scale, lhs, rhs


@pytest.mark.parametrize("log_coef", [False, True])
def test_adjoint_source_accumulators_exact_at_full_geometry(log_coef):
    """K = 64, z 3 -> 0 (3000 steps, 47 checkpoint chunks, per-column EBL
    scales): the source accumulators of the adjoint sweep reproduce the
    forward solve's response to a random injection perturbation to rounding
    error, for every column.  This pins every source-parameter gradient
    without finite differences."""
    prob = AdjointProblem(K=64, z_max=3.0)
    theta = draw_params(64, seed=3)
    solver, plan, st = prob.forward(theta, with_adjoint=True)
    stv = st.detach().clone().requires_grad_(True)
    losses, _ = prob.loss_of_state(stv)
    losses.sum().backward()
    lam_final = stv.grad.permute(0, 2, 1).contiguous()
    ab = solver.adjoint_sweep(plan, lam_final)
    res, lhs, rhs = _linearity_residuals(solver, plan, ab, lam_final, seed=11, log_coef=log_coef)
    print(f"log_coef={log_coef}: max residual {res.max():.3e} (median {np.median(res):.3e})")
    assert np.all(np.isfinite(lhs)) and np.all(np.abs(lhs) > 0)
    assert res.max() <= 1e-10


def test_etd2_propagate_autograd_matches_loss_and_grad():
    """The torch.autograd.Function path: theta (CPU tensor) -> final state ->
    loss -> backward gives the gradient of ``loss_and_grad``."""
    import torch

    from prince_amd.adjoint import ETD2Propagate

    prob = AdjointProblem(K=4, z_max=0.2)
    theta = draw_params(4, seed=17)
    _, grad_ref, _ = prob.loss_and_grad(theta)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    st = ETD2Propagate.apply(th, prob)
    assert st.shape == (prob.run.spec_man.nspec, prob.run.cr_grid.d, 4)
    losses, _ = prob.loss_of_state(st)
    losses.sum().backward()
    g = th.grad.numpy()
    rel = np.abs(g - grad_ref) / np.maximum(np.abs(grad_ref), 1e-300)
    print(f"autograd vs loss_and_grad: max rel {rel.max():.3e}")
    assert th.grad.device.type == "cpu" and np.all(np.isfinite(g))
    assert rel.max() <= 1e-12
