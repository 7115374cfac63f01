"""Result container: behaviour from its specification (log-log interpolation
through positive knots, zero outside their span) and the batched
extraction against the per-column container."""

import numpy as np
import pytest

from prince_amd import workloads
from prince_amd.solvers.result import ResultBatch, UHECRPropagationResult, loglog_onto


def test_loglog_onto_power_laws_and_gaps():
    x = np.logspace(0, 4, 9)
    v = 3.0 * x**-2.7
    v[3] = 0.0  # a dropped knot: interpolation bridges it on the power law
    xt = np.logspace(-1, 5, 61)
    got = loglog_onto(xt, x, v)
    inside = (xt >= x[0]) & (xt <= x[-1])
    assert np.allclose(got[inside], 3.0 * xt[inside] ** -2.7, rtol=1e-12)
    assert not np.any(got[~inside])
    assert not np.any(loglog_onto(xt, x, np.zeros_like(x)))


def test_result_batch_matches_columns():
    run = workloads.nuclear_run()
    rng = np.random.default_rng(4)
    n, K = run.dim_states, 6
    state = np.exp(rng.normal(-5, 3, (n, K)))
    state[rng.random((n, K)) < 0.3] = 0.0  # holes: non-positive knots are skipped
    state[:64, 2] = 0.0  # a species absent in one column
    rb = ResultBatch(state, run.cr_grid.grid, run.spec_man)
    grid, J = rb.group("all", epow=3)
    _, mean, var = rb.lnA("all")
    for k in range(K):
        res = UHECRPropagationResult(state[:, k], run.cr_grid.grid, run.spec_man)
        g1, J1 = res.get_solution_group("all", epow=3)
        _, m1, v1 = res.get_lnA("all")
        assert np.array_equal(g1, grid)
        assert np.allclose(J[:, k], J1, rtol=1e-12, atol=0)
        ok = np.isfinite(m1)
        assert np.array_equal(ok, np.isfinite(mean[:, k]))
        assert np.allclose(mean[ok, k], m1[ok], rtol=1e-12, atol=1e-12)
        assert np.allclose(var[ok, k], v1[ok], rtol=1e-9, atol=1e-10)


def test_result_arithmetic_and_dict_round_trip():
    run = workloads.proton_run()
    s = np.linspace(1.0, 2.0, run.dim_states)
    r = UHECRPropagationResult(s, run.cr_grid.grid, run.spec_man)
    assert np.array_equal((r + r * 2.0).state, 3.0 * s)
    with pytest.raises(TypeError):
        r * np.ones(3)
    other = UHECRPropagationResult(s, run.cr_grid.grid * 2, run.spec_man)
    with pytest.raises(ValueError):
        r + other
    back = UHECRPropagationResult.from_dict(r.to_dict())
    assert np.array_equal(back.state, s) and back.known_species == r.known_species
