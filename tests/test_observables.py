"""The differentiable observables reproduce the result container's accessors."""

import numpy as np
import torch

from prince_amd import workloads
from prince_amd.adjoint import Observables, fractions, make_source, source_derivatives
from prince_amd.solvers.result import UHECRPropagationResult


def _state(run, seed=0):
    rng = np.random.default_rng(seed)
    nspec, d = run.spec_man.nspec, run.cr_grid.d
    E = run.cr_grid.grid
    st = np.zeros((nspec, d))
    for s in range(0, nspec, 3):
        st[s] = rng.uniform(0.5, 2.0) * E ** -rng.uniform(1.5, 3.0)
    st[5, 40:44] = -1e-30  # non-positive knots are dropped
    st[7, :10] = 0.0
    return st


def test_observables_match_result_accessors():
    run = workloads.nuclear_run()
    obs = Observables(run)
    st = _state(run)
    res = UHECRPropagationResult(st.reshape(-1), run.cr_grid.grid, run.spec_man)
    eg, Jref = res.get_solution_group("all", epow=0)
    _, lnA_ref, _ = res.get_lnA("all")
    t = torch.from_numpy(st[:, :, None].copy())
    plan = obs.interp_plan(st[:, :, None])
    l10, lna, J = obs.evaluate(t, plan)
    sel = obs.emask
    Jr = Jref[sel]
    ok = Jr > 0
    assert np.allclose(J.numpy()[:, 0][ok], Jr[ok], rtol=1e-12)
    assert np.all(J.numpy()[:, 0][~ok] == 0)
    lr = lnA_ref[sel]
    assert np.allclose(lna.numpy()[:, 0][ok], lr[ok], rtol=1e-12)


def test_source_parameter_derivatives_by_finite_differences():
    run = workloads.nuclear_run()
    th = np.array([1.2, 9.7, 0.3, -0.5, 1.1, 0.2, 2.5, 1.0])
    g0 = make_source(run, th).injection_grid
    der = source_derivatives(run, th)
    total = np.zeros_like(g0)
    for sidx, g, dgg, dgr, fp in der:
        total[sidx * 64:(sidx + 1) * 64] += g
    assert np.allclose(total, g0, rtol=1e-14)
    for pi, col in ((0, 2), (1, 3)):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[pi] += h
        tm[pi] -= h
        fd = (make_source(run, tp).injection_grid - make_source(run, tm).injection_grid) / (2 * h)
        an = np.zeros_like(g0)
        for d_ in der:
            an[d_[0] * 64:(d_[0] + 1) * 64] += d_[col]
        m = np.abs(an) > 1e-12 * np.abs(an).max()
        assert np.allclose(fd[m], an[m], rtol=1e-6)
    f = fractions(th[2:6])
    assert abs(f.sum() - 1) < 1e-15 and np.all(f > 0)
