"""GPU parity of the uneven batch: every column against its own oracle solve
(its own z_max, hence its own z grid and refresh windows, and its own EBL
scale), <= 1e-8 relative in every bin above 1e-30 of the peak."""

import numpy as np
import pytest

from oracle.etd2_cpu import OracleETD2
from prince_amd import workloads
from prince_amd.solvers.uneven import UnevenBatchSolverETD2, column_field

from .parity import rel_dev

TOL = 1e-8

pytestmark = pytest.mark.gpu


def test_uneven_batch_columns_match_their_own_solves():
    run = workloads.nuclear_run()
    K = 6
    rng = np.random.default_rng(11)
    z_max = rng.uniform(0.1, 0.3, size=K)
    scale = rng.uniform(0.6, 1.4, size=K)
    srcs = workloads.batch_sources(run, K, seed=7)
    g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.solve(dz=1e-3)
    plan = g._plan
    assert plan["nsteps"] == int(np.ceil(z_max.max() / 1e-3 - 1e-9))
    worst = 0.0
    for k in range(K):
        o = OracleETD2(float(z_max[k]), 0.0, run, pfield=column_field(run.photon_field, scale[k]))
        o.add_source_class(srcs[k])
        o.solve(1e-3)
        dev, nb = rel_dev(g.state[:, k], o.state)
        print(f"column {k}: z_max {z_max[k]:.4f} scale {scale[k]:.3f} windows {len(o.refresh_zs)} "
              f"rel dev {dev:.3e} ({nb} bins)")
        assert nb > 100
        worst = max(worst, dev)
    assert worst <= TOL


def test_uneven_batch_equal_columns_equal_plain_solve():
    """All columns with the same z_max and scale 1 reproduce the plain solver."""
    run = workloads.nuclear_run()
    K = 4
    src = workloads.silicon_source(run)
    g = UnevenBatchSolverETD2(np.full(K, 0.12), run)
    for k in range(K):
        g[k].add_source_class(src)
    g.solve(dz=1e-3)
    from prince_amd.solvers.propagation import UHECRPropagationSolverETD2

    p = UHECRPropagationSolverETD2(0.12, 0.0, run)
    p.add_source_class(src)
    p.solve(dz=1e-3)
    for k in range(K):
        dev, nb = rel_dev(g.state[:, k], p.state)
        assert nb > 100 and dev <= TOL, (k, dev)


def test_uneven_chunk_graphs_equal_direct_launches():
    run = workloads.nuclear_run()
    K = 5
    rng = np.random.default_rng(2)
    z_max = rng.uniform(0.1, 0.2, size=K)
    srcs = workloads.batch_sources(run, K, seed=3)
    states = []
    for graphs in (True, False):
        run.backend.use_hip_graphs = graphs
        g = UnevenBatchSolverETD2(z_max, run, ebl_scale=rng.uniform(0.6, 1.4, size=K) * 0 + 1.0)
        for k in range(K):
            g[k].add_source_class(srcs[k])
        g.solve(dz=1e-3)
        assert (g._plan.get("graphs") is not None) == graphs
        states.append(g.state.copy())
    run.backend.use_hip_graphs = True
    assert np.array_equal(states[0], states[1])


def test_uneven_batch_with_em_cascade_matches_column_solves():
    """The uneven batch with the co-evolved EM cascade: every column (own
    z_max, own EBL scale, hence own windows and own device-built cascade
    transfer) against the oracle solve of that column on the field
    CMB + s_k EBL -- nuclear states and photon spectrum."""
    run = workloads.nuclear_em_run()
    K = 4
    rng = np.random.default_rng(13)
    z_max = rng.uniform(0.1, 0.16, size=K)
    scale = rng.uniform(0.6, 1.4, size=K)
    srcs = workloads.batch_sources(run, K, seed=9)
    g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.solve(dz=1e-3)
    n = run.dim_states
    sl = run.em_slices()
    assert g.state.shape == (n + run.dim_em, K)
    base = run.photon_field
    worst_n = worst_g = 0.0
    try:
        for k in range(K):
            run.photon_field = column_field(base, scale[k])
            o = OracleETD2(float(z_max[k]), 0.0, run)
            o.add_source_class(srcs[k])
            o.solve(1e-3)
            dn, nbn = rel_dev(g.state[:n, k], o.state)
            dg, nbg = rel_dev(g.state[sl[22], k], o.em_state[22])
            print(f"column {k}: z_max {z_max[k]:.4f} scale {scale[k]:.3f} nuclear {dn:.3e} "
                  f"({nbn} bins) photons {dg:.3e} ({nbg} bins)")
            assert nbn > 100 and nbg > 10
            worst_n, worst_g = max(worst_n, dn), max(worst_g, dg)
    finally:
        run.photon_field = base
    assert not np.any(g.state[sl[11]]) and not np.any(g.state[sl[-11]])
    assert worst_n <= TOL and worst_g <= TOL


def test_uneven_value_slots_are_bitwise_neutral():
    """Folding each window in place at its first step (one value slot) and
    folding the windows of a step group ahead into two slots give the same
    bits: every column's arithmetic is the same."""
    run = workloads.nuclear_run()
    K = 7
    rng = np.random.default_rng(5)
    z_max = rng.uniform(0.05, 0.16, size=K)
    scale = rng.uniform(0.6, 1.4, size=K)
    srcs = workloads.batch_sources(run, K, seed=9)
    states = {}
    for slots in (1, 2):
        g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
        g.value_slots = slots
        for k in range(K):
            g[k].add_source_class(srcs[k])
        g.solve(dz=1e-3)
        plan = g._plan
        assert plan["nslot"] == slots
        assert (plan["refold_group"] == 1) == (slots == 1)
        live = plan["live"]
        assert live[0] == K and np.all(np.diff(live) <= 0) and live[-1] >= 1
        print(f"slots {slots}: re-fold group {plan['refold_group']} steps, "
              f"live columns {live[0]} -> {live[-1]}")
        states[slots] = g.state.copy()
    assert np.array_equal(states[1], states[2])
