"""GPU: the allocation-free entry point prh_etd2_steps_noalloc against
prh_etd2_steps on the same operator, schedule and inputs (bitwise: the same
kernels in the same order), on two calls in a row, the second with a smaller
input, and its size limit."""

import ctypes

import numpy as np
import pytest

from prince_amd import _lib, workloads
from prince_amd.solvers.propagation import MultiRHSPropagationSolverETD2

pytestmark = pytest.mark.gpu


def _plan(run, K, zi, seed):
    s = MultiRHSPropagationSolverETD2(zi, 0.0, run, K=K)
    srcs = workloads.batch_sources(run, K, seed=seed)
    for k in range(K):
        s[k].add_source_class(srcs[k])
    return s, s.prepare(1e-3)


def _run(s, plan, noalloc):
    import torch

    lib = _lib.load()
    x = plan["x"]
    x.zero_()
    s.fold_windows(plan)
    n = plan["nsteps"]
    if noalloc:
        rc = lib.prh_etd2_steps_noalloc(ctypes.byref(plan["op"]), ctypes.byref(plan["sch"]),
                                        x.data_ptr(), 0, n, plan["stream"])
    else:
        rc = lib.prh_etd2_steps(ctypes.byref(plan["op"]), ctypes.byref(plan["sch"]),
                                ctypes.byref(plan["st"]), 0, n, plan["stream"])
    _lib.check(rc, "steps")
    torch.cuda.synchronize()
    return x.cpu().numpy().copy()


def test_noalloc_matches_existing_entry_point_twice():
    run = workloads.nuclear_run()
    big = _plan(run, 16, 0.08, seed=31)  # dense rows as chunk tasks (dpart in use)
    small = _plan(run, 2, 0.05, seed=32)
    assert big[1]["dense_mode"] == "chunks" and small[1]["dense_mode"] == "chunks"
    for s, plan in (big, small):
        ref = _run(s, plan, noalloc=False)
        got = _run(s, plan, noalloc=True)
        print(f"K={plan['K']}: {np.count_nonzero(ref)} non-zero entries, equal={np.array_equal(ref, got)}")
        assert np.count_nonzero(ref) > 1000
        assert np.array_equal(ref, got)


def test_noalloc_rejects_oversized():
    run = workloads.nuclear_run()
    s, plan = _plan(run, 256, 0.02, seed=33)  # 2 x 40 MB stage vectors > 64 MiB
    lib = _lib.load()
    rc = lib.prh_etd2_steps_noalloc(ctypes.byref(plan["op"]), ctypes.byref(plan["sch"]),
                                    plan["x"].data_ptr(), 0, plan["nsteps"], plan["stream"])
    assert rc != 0
    msg = lib.prh_last_error().decode()
    assert "bytes of scratch" in msg and str(_lib.SCRATCH_BYTES) in msg
