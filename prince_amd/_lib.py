"""ctypes binding of ``libprince_hip.so`` (C ABI declared in include/prince_hip.h).

The library is built in-tree next to this file (``prince_amd/build.py``).
Loading fails loudly when it is missing: there is no fallback path.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libprince_hip.so")

EXPORTED = (
    "prh_version",
    "prh_last_error",
    "prh_fold_gemm",
    "prh_gather_rows",
    "prh_etd2_steps",
    "prh_etd2_steps_noalloc",
    "prh_etd2_stage",
    "prh_graph_create",
    "prh_graph_run",
    "prh_graph_destroy",
    "prh_adjoint_step",
    "prh_adjoint_seed",
    "prh_adjoint_steps",
    "prh_em_step",
    "prh_field_tables",
    "prh_fold_rows",
    "prh_uneven_step",
    "prh_uneven_graph_create",
    "prh_cascade_transfer",
    "prh_cascade_bh",
    "prh_toeplitz_rows",
    "prh_em_transfer_stack",
    "prh_em_step_transfer",
)

WG_WAVES = 16
#: size of the static device scratch of prh_etd2_steps_noalloc (PRH_SCRATCH_BYTES)
SCRATCH_BYTES = 64 << 20

PRH_F_PARTIAL_DIFF = 1
PRH_F_INJECTION = 2
PRH_F_ROW_PAIRS = 4
PRH_TASK_KIND_MASK = 3
PRH_TASK_TO_DPART = 4
PRH_TASK_PAIR = 8
PRH_TASK_PAIR_SHIFT = 8

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class PrhOperator(ctypes.Structure):
    _fields_ = [
        ("nspec", _i32),
        ("d", _i32),
        ("K", _i32),
        ("ntasks", _i32),
        ("tasks", _p),
        ("blocks", _p),
        ("mvals_stride", _i64),
        ("diag_off", _i64),
        ("mvals", _p),
        ("egrid", _p),
        ("dcoef", _p),
        ("ddiag", _p),
        ("pscale", _p),
        ("rpair", _p),
        ("src_slot", _p),
        ("nsrc", _i32),
        ("flags", _i32),
        ("src_grid", _p),
        ("kt", _i32),
        ("pad_", _i32),
        ("mvals_e", _p),
        ("rpair_e", _p),
        ("escale", _p),
        ("wg_waves", _i32),
        ("ndense", _i32),
        ("dchunk", _i32),
        ("ndchunk", _i32),
        ("dense_row", _p),
        ("dense_tab", _p),
        ("nfin", _i32),
        ("pad3_", _i32),
        ("fin_tasks", _p),
    ]


class PrhSchedule(ctypes.Structure):
    _fields_ = [
        ("nsteps", _i32),
        ("pad_", _i32),
        ("h", _p),
        ("dldz", _p),
        ("hadia", _p),
        ("finj", _p),
        ("win", _p),
        ("src_coef", _p),
    ]


class PrhState(ctypes.Structure):
    _fields_ = [
        ("x", _p),
        ("a", _p),
        ("fphi", _p),
        ("step_ctl", _p),
        ("rec_x", _p),
        ("rec_a", _p),
        ("rec_fx", _p),
        ("rec_fa", _p),
        ("rec_first", _i32),
        ("rec_count", _i32),
        ("em", _p),
        ("dpart", _p),
    ]


class PrhEm(ctypes.Structure):
    _fields_ = [
        ("d_em", _i32),
        ("nwin", _i32),
        ("egrid", _p),
        ("dcoef", _p),
        ("ddiag", _p),
        ("cbh", _p),
        ("bhw", _p),
        ("rbh", _p),
        ("tg", _p),
        ("te", _p),
        ("gam", _p),
        ("wbh", _p),
        ("gam2", _p),
        ("tstack", _p),
        ("qstep", _p),
    ]


EM_MAX = 144


class PrhUneven(ctypes.Structure):
    _fields_ = [
        ("nsteps", _i32),
        ("nwin", _i32),
        ("h", _p),
        ("dldz", _p),
        ("hadia", _p),
        ("finj", _p),
        ("win", _p),
        ("src_coef", _p),
        ("rpair", _p),
        ("mrhs", _p),
        ("nslot", _i32),
        ("pad_", _i32),
        ("live", _p),
    ]


class PrhRefold(ctypes.Structure):
    _fields_ = [
        ("kern", _p),
        ("nrows", _i64),
        ("dph", _i32),
        ("enabled", _i32),
        ("phot", _p),
        ("in_rows", _p),
        ("out_rows", _p),
        ("ref_off", _p),
    ]


class PrhAdjoint(ctypes.Structure):
    _fields_ = [
        ("lam", _p),
        ("lam_a", _p),
        ("u", _p),
        ("v", _p),
        ("rec_x", _p),
        ("rec_a", _p),
        ("rec_fx", _p),
        ("rec_fa", _p),
        ("rec_first", _i32),
        ("rec_count", _i32),
        ("gsrc0", _p),
        ("gsrc1", _p),
        ("lnz1", _p),
        ("gdot", _p),
    ]


_LIB = None


class HipLibraryError(RuntimeError):
    pass


class PrhToeplitz(ctypes.Structure):
    _fields_ = [
        ("d", _i32),
        ("dph", _i32),
        ("n1", _i32),
        ("n2", _i32),
        ("nrows", _i64),
        ("desc", _p),
        ("tab1d", _p),
        ("tab2d", _p),
        ("fmo", _p),
        ("fdiff", _p),
    ]


class PrhCascade(ctypes.Structure):
    _fields_ = [
        ("n", _i32),
        ("d", _i32),
        ("nb", _i32),
        ("max_scatter", _i32),
        ("tol", ctypes.c_double),
        ("cool_floor", ctypes.c_double),
        ("bh_emin", ctypes.c_double),
        ("E", _p),
        ("dE", _p),
        ("pref_emis", _p),
        ("pref_pair", _p),
        ("cic_b", _p),
        ("cic_f", _p),
        ("Ecr", _p),
        ("eps_b", _p),
        ("kern_b", _p),
    ]


def load():
    """Load (once) and return the ctypes library handle."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise HipLibraryError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m prince_amd.build`; the HIP backend has no fallback."
        )
    # torch ships its own HIP runtime; loading it first makes this library bind
    # to that same runtime (one HIP context per process, so the stream handles
    # and device pointers torch hands over are valid here)
    import torch  # noqa: F401

    lib = ctypes.CDLL(LIB_PATH)
    lib.prh_version.restype = _i32
    lib.prh_last_error.restype = ctypes.c_char_p
    lib.prh_fold_gemm.restype = _i32
    lib.prh_fold_gemm.argtypes = [_p, _i64, _i32, _p, _i32, _p, _i64, _p]
    lib.prh_gather_rows.restype = _i32
    lib.prh_gather_rows.argtypes = [_p, _p, _i64, _i32, _p, _p]
    lib.prh_etd2_steps.restype = _i32
    lib.prh_etd2_steps.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        ctypes.POINTER(PrhState),
        _i32,
        _i32,
        _p,
    ]
    lib.prh_etd2_steps_noalloc.restype = _i32
    lib.prh_etd2_steps_noalloc.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        _p,
        _i32,
        _i32,
        _p,
    ]
    lib.prh_cascade_transfer.restype = _i32
    lib.prh_cascade_transfer.argtypes = [ctypes.POINTER(PrhCascade), _p, _i64, _p, _p, _i32, _p,
                                         _i64, _p, _p, _p, _p]
    lib.prh_toeplitz_rows.restype = _i32
    lib.prh_toeplitz_rows.argtypes = [ctypes.POINTER(PrhToeplitz), _p, _i64, _p, _p]
    lib.prh_em_transfer_stack.restype = _i32
    lib.prh_em_transfer_stack.argtypes = [_p, _p, _p, _i32, _i32, _i32, _p, _p]
    lib.prh_em_step_transfer.restype = _i32
    lib.prh_em_step_transfer.argtypes = [
        ctypes.POINTER(PrhOperator), ctypes.POINTER(PrhSchedule), ctypes.POINTER(PrhEm), _p, _p]
    lib.prh_cascade_bh.restype = _i32
    lib.prh_cascade_bh.argtypes = [ctypes.POINTER(PrhCascade), _p, _i32, _p, _p]
    lib.prh_etd2_stage.restype = _i32
    lib.prh_etd2_stage.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        ctypes.POINTER(PrhState),
        _i32,
        _i32,
        _p,
    ]
    _d = ctypes.c_double
    lib.prh_field_tables.restype = _i32
    lib.prh_field_tables.argtypes = [_p, _p, _i32, _p, _p, _p, _i32, _i32, _d, _d, _d, _p, _p]
    lib.prh_fold_rows.restype = _i32
    lib.prh_fold_rows.argtypes = [_p, _i64, _i32, _p, _p, _p, _i32, _p, _i64, _p]
    lib.prh_uneven_step.restype = _i32
    lib.prh_uneven_step.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhUneven),
        ctypes.POINTER(PrhState),
        _i32,
        _p,
    ]
    lib.prh_uneven_graph_create.restype = _i32
    lib.prh_uneven_graph_create.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhUneven),
        ctypes.POINTER(PrhState),
        ctypes.POINTER(PrhRefold),
        _i32,
        _i32,
        ctypes.POINTER(_p),
    ]
    lib.prh_em_step.restype = _i32
    lib.prh_em_step.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        ctypes.POINTER(PrhState),
        _i32,
        _p,
    ]
    for name in ("prh_adjoint_step", "prh_adjoint_seed"):
        fn = getattr(lib, name)
        fn.restype = _i32
        fn.argtypes = [
            ctypes.POINTER(PrhOperator),
            ctypes.POINTER(PrhSchedule),
            ctypes.POINTER(PrhAdjoint),
            _i32,
            _p,
        ]
    lib.prh_adjoint_steps.restype = _i32
    lib.prh_adjoint_steps.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        ctypes.POINTER(PrhAdjoint),
        _i32,
        _i32,
        _p,
    ]
    lib.prh_graph_create.restype = _i32
    lib.prh_graph_create.argtypes = [
        ctypes.POINTER(PrhOperator),
        ctypes.POINTER(PrhSchedule),
        ctypes.POINTER(PrhState),
        _i32,
        ctypes.POINTER(_p),
    ]
    lib.prh_graph_run.restype = _i32
    lib.prh_graph_run.argtypes = [_p, _i32, _i32, _p]
    lib.prh_graph_destroy.restype = _i32
    lib.prh_graph_destroy.argtypes = [_p]
    _LIB = lib
    return lib


class StepGraph:
    """Owned handle of a captured step graph (destroyed with the object)."""

    def __init__(self, op, sch, st, chunk):
        lib = load()
        h = _p()
        check(
            lib.prh_graph_create(ctypes.byref(op), ctypes.byref(sch), ctypes.byref(st), int(chunk),
                                 ctypes.byref(h)),
            "step graph capture",
        )
        self._h = h
        self.chunk = int(chunk)

    def run(self, k_begin, k_end, stream):
        check(load().prh_graph_run(self._h, int(k_begin), int(k_end), stream), "step graph run")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().prh_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UnevenChunkGraph(StepGraph):
    """Owned handle of a captured uneven-batch chunk (steps baked in)."""

    def __init__(self, op, un, st, rf, k_begin, k_end):
        lib = load()
        h = _p()
        check(
            lib.prh_uneven_graph_create(ctypes.byref(op), ctypes.byref(un), ctypes.byref(st),
                                        ctypes.byref(rf), int(k_begin), int(k_end), ctypes.byref(h)),
            "uneven chunk capture",
        )
        self._h = h
        self.chunk = int(k_end - k_begin)
        self.k_begin, self.k_end = int(k_begin), int(k_end)

    def launch(self, stream):
        self.run(self.k_begin, self.k_end, stream)


def check(status, what):
    if status != 0:
        msg = load().prh_last_error().decode(errors="replace")
        raise HipLibraryError(f"{what} failed: {msg}")
