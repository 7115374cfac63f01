"""GPU parity at the full benchmark sizes: the A <= 56 plain solve and columns
of the K = 512 batch at z 3 -> 0 (3000 steps, 288 windows) against oracle
fixtures, the narrow-layout nuclear EM cascade at K = 16 against the oracle,
and the device-derived off-diagonal pattern against the oracle's
``split_operator`` pattern, index for index."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.etd2_cpu import OracleETD2, split_operator
from prince_amd import workloads
from prince_amd.solvers.propagation import (
    MultiRHSPropagationSolverETD2,
    UHECRPropagationSolverETD2,
)

from .parity import rel_dev
from .test_golden import BATCH, batch_fixture_sources, load

TOL = 1e-8

pytestmark = pytest.mark.gpu

#: the stage configuration the benchmark's autotuner picks at K = 512 (8
#: right-hand sides per wave, part budget 48, 8-wave workgroups, dense rows as
#: ordinary split rows, row-pair tasks)
BENCH_PLAN_512 = {("kt", 512): 8, ("part_cost", 512): 48, ("wg_waves", 512): 8,
                  ("dense", 512): False, ("pairs", 512): True}


def test_plain_nuclear_z3_against_fixture():
    g = load("etd2_nuclear_si28_z3.npz")
    run = workloads.nuclear_run()
    s = UHECRPropagationSolverETD2(3.0, 0.0, run)
    s.add_source_class(workloads.silicon_source(run))
    s.solve(dz=1e-3)
    assert len(s.refresh_zs) == 288 and np.array_equal(s.refresh_zs, g["refresh_zs"])
    assert s._plan["dense_mode"] == "chunks"
    dev, nb = rel_dev(s.state, g["state"])
    print(f"plain A<=56 z 3->0: {nb} bins, max rel dev {dev:.3e}")
    assert np.all(np.isfinite(s.state))
    assert nb > 1000 and dev <= TOL


def test_batch512_bench_plan_against_fixture_columns():
    """The K = 512 batched solve with the benchmark's stage configuration:
    the fixture's columns (0, 1, 2, 3, 511) against their oracle solves."""
    g = load(BATCH)
    run = workloads.nuclear_run()
    K = int(g["K"])
    srcs = workloads.batch_sources(run, K, seed=int(g["seed"]))
    s = MultiRHSPropagationSolverETD2(3.0, 0.0, run, K=K)
    s._tuned = dict(BENCH_PLAN_512)
    for k in range(K):
        s[k].add_source_class(srcs[k])
    s.solve(dz=1e-3)
    plan = s._plan
    assert plan["kt"] == 8 and plan["dense_mode"] is None
    assert plan["op"].wg_waves == 8 and plan["nwin"] == 288
    assert int(np.sum((plan["tasks"][:, 5] & 8) != 0)) > 100  # row pairs
    worst = 0.0
    for i, k in enumerate(g["columns"]):
        dev, nb = rel_dev(s.state[:, k], g["state"][:, i])
        print(f"column {k}: {nb} bins, max rel dev {dev:.3e}")
        assert nb > 1000 and dev <= TOL, (k, dev)
        worst = max(worst, dev)
    # the rebuilt fixture sources are the same objects' draws
    fs = batch_fixture_sources(run, g)
    assert np.array_equal(fs[0].injection_grid, srcs[int(g["columns"][0])].injection_grid)
    print(f"K=512 bench plan: worst column {worst:.3e}")


def test_em_cascade_nuclear_narrow_k16():
    """Narrow layout (lanes = bins) at K = 16: the Bethe-Heitler weights run
    one workgroup per right-hand side over all ~300 charged species in
    8-species batches; nuclear and photon states of every column against
    the oracle."""
    run = workloads.nuclear_em_run()
    K = 16
    srcs = workloads.batch_sources(run, K, seed=15)
    g = MultiRHSPropagationSolverETD2(0.1, 0.0, run, K=K)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.solve(dz=1e-3)
    assert g._plan["em"] is not None
    o = OracleETD2(0.1, 0.0, run, K=K)
    for k in range(K):
        o.add_source_class(srcs[k], k)
    o.solve(1e-3)
    n = run.dim_states
    sl = run.em_slices()
    worst_n = worst_g = 0.0
    for k in range(K):
        dn, nbn = rel_dev(g.state[:n, k], o.state[:, k])
        dg, nbg = rel_dev(g.state[sl[22], k], o.em_state[22][:, k])
        assert nbn > 100 and nbg > 10
        worst_n, worst_g = max(worst_n, dn), max(worst_g, dg)
    assert not np.any(g.state[sl[11]]) and not np.any(g.state[sl[-11]])
    print(f"nuclear+EM narrow K=16: nuclear {worst_n:.3e}, photons {worst_g:.3e}")
    assert worst_n <= TOL and worst_g <= TOL


@pytest.mark.parametrize("photohad", [True, False])
def test_off_diagonal_pattern_equals_oracle_split(photohad):
    """The off-diagonal pattern the device derives from its first-window fold
    is the oracle's ``split_operator`` pattern of the host fold, entry for
    entry (reference etd2.py:234-268, propagation.py:721-807)."""
    run = workloads.nuclear_run()
    s = UHECRPropagationSolverETD2(3.0, 0.0, run, enable_photohad_losses=photohad)
    s.add_source_class(workloads.silicon_source(run))
    plan = s.prepare(1e-3)
    lay = plan["layout"].fwd
    ir = run.int_rates
    M = ir.coupling_mat.copy()
    M.data = ir._batch_matrix @ ir.photon_vector(float(plan["zs"][0]))
    if not photohad:
        M.data[:] = 0.0
    _, M_off = split_operator(M)
    M_off.sort_indices()
    n = run.dim_states
    rows = np.repeat(np.arange(n), np.diff(ir.coupling_mat.indptr))
    dev_rows = rows[lay.off_to_M]
    dev_cols = ir.coupling_mat.indices[lay.off_to_M]
    ref_rows = np.repeat(np.arange(n), np.diff(M_off.indptr))
    assert lay.nnz_off == M_off.nnz
    assert np.array_equal(dev_rows, ref_rows)
    assert np.array_equal(dev_cols, M_off.indices)
    # every kept entry has its own slot of the block-band layout
    assert np.unique(lay.off_slot).size == lay.off_slot.size
    assert lay.off_slot.size == 0 or lay.off_slot.max() < lay.zslots
    # the device's window-0 values equal the host fold: off-diagonal entries at
    # their block-band slots, the diagonal in the window's first n values
    wl = plan["layout"]
    s.fold_windows(plan)
    vals = plan["bufs"]["mvals"][0].cpu().numpy()
    if M_off.nnz:
        np.testing.assert_allclose(vals[wl.fwd_off + lay.off_slot], M.data[lay.off_to_M],
                                   rtol=1e-12, atol=0)
    diag = vals[wl.diag_off:wl.diag_off + n]
    ref_diag = M.diagonal()
    np.testing.assert_allclose(diag, ref_diag, rtol=1e-12, atol=1e-12 * np.abs(ref_diag).max())
    assert isinstance(M_off, sp.csr_matrix)
    print(f"photohad={photohad}: {M_off.nnz} off-diagonal entries equal")
    s.close()


def test_uneven_em_columns_against_oracle_over_many_windows():
    """The uneven batch with the EM cascade over many refresh windows: columns
    from z_max 1.3 and 0.7 (130 and 70 windows, each with its own EBL scale and
    its own device-built cascade transfers) against oracle solves on CMB + s EBL;
    two shorter columns retire early.  Nuclear states and photons <= 1e-8."""
    from prince_amd.solvers.uneven import UnevenBatchSolverETD2, column_field

    run = workloads.nuclear_em_run()
    z_max = np.array([1.3, 0.7, 0.25, 0.12])
    scale = np.array([0.7, 1.3, 1.1, 0.9])
    K = z_max.size
    srcs = workloads.batch_sources(run, K, seed=21)
    g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    g.solve(dz=1e-3)
    n = run.dim_states
    sl = run.em_slices()
    base = run.photon_field
    try:
        for k in (0, 1):
            run.photon_field = column_field(base, scale[k])
            o = OracleETD2(float(z_max[k]), 0.0, run)
            o.add_source_class(srcs[k])
            o.solve(1e-3)
            dn, nbn = rel_dev(g.state[:n, k], o.state)
            dg, nbg = rel_dev(g.state[sl[22], k], o.em_state[22])
            print(f"column {k}: z_max {z_max[k]} ({len(o.refresh_zs)} windows) scale {scale[k]}: "
                  f"nuclear {dn:.3e} ({nbn} bins), photons {dg:.3e} ({nbg} bins)")
            assert nbn > 1000 and nbg > 50
            assert dn <= TOL and dg <= TOL
    finally:
        run.photon_field = base
