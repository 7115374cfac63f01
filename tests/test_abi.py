"""C ABI: the in-tree library loads and exports every function the header declares;
host-side guards of the HIP backend (no GPU needed)."""

import os
import re

import pytest

from prince_amd import _lib
from prince_amd.model.config import BackendConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\s*(?:int32_t|const char\*)\s+(prh_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_declared_symbols():
    lib = _lib.load()
    declared = _declared()
    assert declared, "no declarations found"
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.prh_version() >= 2
    assert isinstance(lib.prh_last_error(), bytes)


def test_abi_rejects_bad_arguments_without_touching_the_gpu():
    lib = _lib.load()
    assert lib.prh_fold_gemm(None, 10, 5000, None, 1, None, 10, None) != 0
    assert b"dph" in lib.prh_last_error()
    assert lib.prh_gather_rows(None, None, -1, 72, None, None) != 0
    assert lib.prh_etd2_steps(None, None, None, 0, 1, None) != 0
    assert b"null" in lib.prh_last_error()
    assert lib.prh_graph_run(None, 0, 1, None) != 0
    assert lib.prh_graph_destroy(None) == 0


def test_backend_dispatch_and_gpu_requirement():
    import torch

    from prince_amd import workloads
    from prince_amd.solvers.propagation import UHECRPropagationSolverETD2

    run = workloads.proton_run()
    saved = run.backend
    try:
        run.backend = BackendConfig(linear_algebra_backend="scipy")
        with pytest.raises(NotImplementedError):
            UHECRPropagationSolverETD2(1.0, 0.0, run)
    finally:
        run.backend = saved
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="GPU"):
            UHECRPropagationSolverETD2(1.0, 0.0, run)


def test_abi_rejects_bad_operators_without_touching_the_gpu():
    """Operator preconditions are checked before any launch."""
    import ctypes

    lib = _lib.load()
    dummy = ctypes.c_int32(0)
    ptr = ctypes.addressof(dummy)
    op = _lib.PrhOperator(nspec=2, d=64, K=96, ntasks=0, kt=1)
    sch = _lib.PrhSchedule(nsteps=4)
    st = _lib.PrhState(x=ptr, a=ptr, fphi=ptr, step_ctl=ptr)
    args = (ctypes.byref(op), ctypes.byref(sch), ctypes.byref(st), 0, 1, None)
    assert lib.prh_etd2_steps(*args) != 0
    assert b"empty operator" in lib.prh_last_error()
    op.ntasks = 16
    op.kt = 3
    assert lib.prh_etd2_steps(*args) != 0
    assert b"kt must be" in lib.prh_last_error()
    op.kt = 1
    op.wg_waves = 4
    assert lib.prh_etd2_steps(*args) != 0
    assert b"wg_waves must be" in lib.prh_last_error()
    op.wg_waves = 8
    adj = _lib.PrhAdjoint(lam=ptr, lam_a=ptr, u=ptr, v=ptr)
    assert lib.prh_adjoint_step(ctypes.byref(op), ctypes.byref(sch), ctypes.byref(adj), 0, None) != 0
    assert b"16-wave" in lib.prh_last_error()


def test_row_pair_tasks_are_refused_where_no_kernel_runs_them():
    """An operator flagged PRH_F_ROW_PAIRS is refused on the host by the
    two-field stages, the adjoint and the uneven batch (their kernels would run
    a pair task as one long row)."""
    import ctypes

    lib = _lib.load()
    dummy = ctypes.c_int32(0)
    ptr = ctypes.addressof(dummy)
    op = _lib.PrhOperator(nspec=2, d=64, K=96, ntasks=16, kt=1, flags=_lib.PRH_F_ROW_PAIRS,
                          mvals_e=ptr, escale=ptr, rpair_e=ptr)
    sch = _lib.PrhSchedule(nsteps=4)
    st = _lib.PrhState(x=ptr, a=ptr, fphi=ptr, step_ctl=ptr)
    assert lib.prh_etd2_steps(ctypes.byref(op), ctypes.byref(sch), ctypes.byref(st), 0, 1, None) != 0
    assert b"row-pair" in lib.prh_last_error()
    op.mvals_e = op.escale = op.rpair_e = None
    adj = _lib.PrhAdjoint(lam=ptr, lam_a=ptr, u=ptr, v=ptr)
    assert lib.prh_adjoint_step(ctypes.byref(op), ctypes.byref(sch), ctypes.byref(adj), 0, None) != 0
    assert b"row-pair" in lib.prh_last_error()
    op.kt = 2
    op.mvals_stride = 128
    op.tasks = ptr
    un = _lib.PrhUneven(nsteps=2, nwin=2, h=ptr, dldz=ptr, hadia=ptr, finj=ptr, win=ptr,
                        rpair=ptr, mrhs=ptr, nslot=2)
    assert lib.prh_uneven_step(ctypes.byref(op), ctypes.byref(un), ctypes.byref(st), 0, None) != 0
    assert b"row-pair" in lib.prh_last_error()


def test_noalloc_entry_point_rejects_before_touching_the_gpu():
    """Null arguments and a scratch need beyond the static array are refused on
    the host, before any device call."""
    import ctypes

    lib = _lib.load()
    op = _lib.PrhOperator(nspec=305, d=64, K=256)
    sch = _lib.PrhSchedule(nsteps=1)
    assert lib.prh_etd2_steps_noalloc(None, ctypes.byref(sch), None, 0, 1, None) != 0
    assert b"null" in lib.prh_last_error()
    rc = lib.prh_etd2_steps_noalloc(ctypes.byref(op), ctypes.byref(sch), 8, 0, 1, None)
    assert rc != 0 and b"bytes of scratch" in lib.prh_last_error()


def test_uneven_step_rejects_bad_slots_and_live_counts_on_the_host():
    """prh_uneven_step refuses a value-slot count other than 1 or 2 and a live
    column count outside [0, K] before any launch."""
    import ctypes

    lib = _lib.load()
    dummy = ctypes.c_int32(0)
    ptr = ctypes.addressof(dummy)
    op = _lib.PrhOperator(nspec=2, d=64, K=4, ntasks=16, tasks=ptr, kt=2, mvals_stride=128)
    st = _lib.PrhState(x=ptr, a=ptr, fphi=ptr, step_ctl=ptr)
    live = (ctypes.c_int32 * 2)(4, 5)
    un = _lib.PrhUneven(nsteps=2, nwin=2, h=ptr, dldz=ptr, hadia=ptr, finj=ptr, win=ptr,
                        rpair=ptr, mrhs=ptr, nslot=3, live=ctypes.addressof(live))
    args = (ctypes.byref(op), ctypes.byref(un), ctypes.byref(st))
    assert lib.prh_uneven_step(*args, 0, None) != 0
    assert b"nslot" in lib.prh_last_error()
    un.nslot = 2
    assert lib.prh_uneven_step(*args, 1, None) != 0
    assert b"live column count" in lib.prh_last_error()
