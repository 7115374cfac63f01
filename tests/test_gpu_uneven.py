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
