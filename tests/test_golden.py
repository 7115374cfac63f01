"""Golden fixtures (tests/golden/, written by make_golden.py from the FP64
oracle): the oracle still reproduces them, and the files hold what the GPU
tests compare against."""

import os

import numpy as np
import pytest

from oracle.etd2_cpu import OracleETD2
from prince_amd import workloads

from .parity import rel_dev

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [
    ("etd2_proton_cmb_z0.05.npz", workloads.proton_run, workloads.proton_source),
    ("etd2_nuclear_si28_z0.1.npz", workloads.nuclear_run, workloads.silicon_source),
    ("etd2_nuclear_si28_z3.npz", workloads.nuclear_run, workloads.silicon_source),
    ("etd2_proton_em_z0.05.npz", workloads.proton_em_run, workloads.proton_source),
]

BATCH = "etd2_nuclear_batch512_z3_cols.npz"


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def batch_fixture_sources(run, g):
    """The source objects of the fixture's columns, rebuilt from its stored draws."""
    return [workloads.batch_source(run, g["gamma"][i], g["lrmax"][i], g["frac"][i], g["m"][i])
            for i in range(len(g["columns"]))]


@pytest.mark.parametrize("name,run_fn,src_fn", CASES)
def test_oracle_reproduces_golden(name, run_fn, src_fn):
    g = load(name)
    run = run_fn()
    assert np.array_equal(g["egrid"], run.cr_grid.grid)
    assert np.array_equal(g["pdg"], np.array(run.spec_man.known_species))
    o = OracleETD2(float(g["initial_z"]), float(g["final_z"]), run)
    o.add_source_class(src_fn(run))
    o.solve(float(g["dz"]))
    assert np.allclose(o.refresh_zs, g["refresh_zs"], rtol=0, atol=0)
    dev, nb = rel_dev(o.state, g["state"])
    assert nb > 30 and dev < 1e-10
    if "photons" in g:
        dev, nb = rel_dev(o.em_state[22], g["photons"])
        assert nb > 30 and dev < 1e-10


def test_batch_fixture_draws_and_oracle():
    """The batch fixture's stored draws are the seeded draws of those columns,
    and the oracle reproduces every column (z 3 -> 0, 288 windows)."""
    g = load(BATCH)
    run = workloads.nuclear_run()
    K, seed, cols = int(g["K"]), int(g["seed"]), g["columns"]
    gam, lr, fr, m = workloads.batch_draws(K, seed)
    assert np.array_equal(g["gamma"], gam[cols]) and np.array_equal(g["lrmax"], lr[cols])
    assert np.array_equal(g["frac"], fr[cols]) and np.array_equal(g["m"], m[cols])
    assert len(g["refresh_zs"]) == 288
    srcs = batch_fixture_sources(run, g)
    o = OracleETD2(float(g["initial_z"]), float(g["final_z"]), run, K=len(cols))
    for i, s in enumerate(srcs):
        o.add_source_class(s, i)
    o.solve(float(g["dz"]))
    for i in range(len(cols)):
        dev, nb = rel_dev(o.state[:, i], g["state"][:, i])
        assert nb > 1000 and dev < 1e-10, (i, dev)
