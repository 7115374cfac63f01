"""GPU parity of the co-evolved EM cascade: nuclear states and the photon
spectrum against the CPU oracle (<= 1e-8 relative in every bin above 1e-30 of
each block's peak), for one column (K = 1) and for K = 64."""

import numpy as np
import pytest

from oracle.etd2_cpu import OracleETD2
from prince_amd import workloads
from prince_amd.solvers.propagation import (
    MultiRHSPropagationSolverETD2,
    UHECRPropagationSolverETD2,
)

from .parity import rel_dev

TOL = 1e-8

pytestmark = pytest.mark.gpu


def _split(run, state):
    n = run.dim_states
    sl = run.em_slices()
    return state[:n], state[sl[22]], state[sl[11]], state[sl[-11]]


@pytest.mark.parametrize("zi", [0.05, 0.051])  # even and odd step counts (photon buffers alternate)
def test_em_cascade_proton_single(zi):
    run = workloads.proton_em_run()
    src = workloads.proton_source(run)
    g = UHECRPropagationSolverETD2(zi, 0.0, run)
    g.add_source_class(src)
    g.solve(dz=1e-3)
    o = OracleETD2(zi, 0.0, run)
    o.add_source_class(src)
    o.solve(1e-3)
    nuc, gam, em_, ep_ = _split(run, g.state)
    onuc, ogam = o.state, o.em_state[22]
    dn, nbn = rel_dev(nuc, onuc)
    dg, nbg = rel_dev(gam, ogam)
    print(f"proton+EM: nuclear {dn:.3e} ({nbn} bins), photons {dg:.3e} ({nbg} bins)")
    assert nbn > 30 and nbg > 30
    assert dn <= TOL and dg <= TOL
    assert not np.any(em_) and not np.any(ep_)
    e, spec = g.res.get_solution(22)
    assert e.size == run.em_grid.d and np.array_equal(spec, gam)


@pytest.mark.parametrize("dense", ["chunks", False])
def test_em_cascade_nuclear_columns(dense):
    """z 0.1 -> 0.  (After only ~30 steps a few bins hold negative undershoots
    of ~1e-28 of the peak whose value is set by the last bit of exp(h d); a
    1-ulp change of exp on the CPU moves them by up to 4e-6, so short runs are
    not a meaningful 1e-8 comparison.)"""
    run = workloads.nuclear_em_run()
    K = 64
    srcs = workloads.batch_sources(run, K, seed=5)
    g = MultiRHSPropagationSolverETD2(0.1, 0.0, run, K=K)
    g._tuned = {("dense", K): dense}
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.solve(dz=1e-3)
    assert g._plan["dense_mode"] == (dense or None)
    o = OracleETD2(0.1, 0.0, run, K=K)
    for k in range(K):
        o.add_source_class(srcs[k], k)
    o.solve(1e-3)
    nuc, gam, _, _ = _split(run, g.state)
    worst_n = worst_g = 0.0
    for k in range(K):
        dn, nbn = rel_dev(nuc[:, k], o.state[:, k])
        dg, nbg = rel_dev(gam[:, k], o.em_state[22][:, k])
        assert nbn > 100 and nbg > 10
        worst_n, worst_g = max(worst_n, dn), max(worst_g, dg)
    print(f"nuclear+EM K=64 dense={dense}: nuclear {worst_n:.3e}, photons {worst_g:.3e}")
    assert worst_n <= TOL and worst_g <= TOL


def test_em_cascade_graph_equals_direct():
    run = workloads.proton_em_run()
    states = []
    for graphs in (True, False):
        run.backend.use_hip_graphs = graphs
        g = UHECRPropagationSolverETD2(0.08, 0.0, run)
        g.add_source_class(workloads.proton_source(run))
        g.solve(dz=1e-3)
        states.append(g.state.copy())
    run.backend.use_hip_graphs = True
    assert np.array_equal(states[0], states[1])


def test_em_execute_restores_initial_photons():
    """prepare() + execute() twice: each run starts from the plan's initial
    photon state (gam0), so both give the same photon and nuclear states."""
    import torch

    run = workloads.proton_em_run()
    g = UHECRPropagationSolverETD2(0.07, 0.0, run)
    g.add_source_class(workloads.proton_source(run))
    plan = g.prepare(1e-3)
    eb = plan["em_bufs"]
    eb["gam0"].copy_(torch.linspace(1.0, 2.0, eb["gam0"].numel(), dtype=torch.float64)
                     .reshape(eb["gam0"].shape))
    outs = []
    for _ in range(2):
        x = g.execute(plan)
        torch.cuda.synchronize()
        outs.append((x.cpu().numpy().copy(), g.em_host_state(plan).copy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.any(outs[0][1] != 0.0)


def test_em_step_transfer_is_T_gamma_times_the_oracle_photon_step():
    """qstep[t] = T_gamma(w_t) P(t), where P(t) is the oracle's photon ETD2 step
    (etd2_step with the adiabatic FD operator, etd2.py:412-439) applied to the
    unit vectors, at the first, second, middle and last steps of a solve."""
    import torch

    run = workloads.proton_em_run()
    g = UHECRPropagationSolverETD2(0.05, 0.0, run)
    g.add_source_class(workloads.proton_source(run))
    plan = g.prepare(1e-3)
    torch.cuda.synchronize()
    Q = plan["em_bufs"]["qstep"].cpu().numpy()
    Tg = plan["em_bufs"]["tg"].cpu().numpy()
    zg, win = plan["z_grid"], np.asarray(plan["win"])
    de = run.em_grid.d
    o = OracleETD2(0.05, 0.0, run)
    nst = len(zg) - 1
    assert Q.shape == (nst, de, de)
    worst = worst_step = 0.0
    for t in (0, 1, nst // 2, nst - 1):
        em = {22: np.eye(de)}
        o._em_step(zg[t], zg[t + 1] - zg[t], em)  # column j: the step of unit vector j
        T = Tg[win[t]]
        ref = T @ em[22]
        dev = float(np.max(np.abs(Q[t] - ref)) / np.max(np.abs(ref)))
        # the step's own contribution T (P - I), where a dropped or wrong term shows
        step = float(np.max(np.abs((Q[t] - T) - (ref - T))) / np.max(np.abs(ref - T)))
        worst, worst_step = max(worst, dev), max(worst_step, step)
    print(f"T_gamma P(t) against T_gamma times the oracle's photon step: {worst:.2e} "
          f"(step part {worst_step:.2e})")
    assert worst < 1e-12 and worst_step < 1e-9
