"""ETD2 propagation solvers with the MI355X HIP backend.

Public surface mirrors the reference (``solvers/propagation.py``):

* ``UHECRPropagationSolverETD2(initial_z, final_z, prince_run, enable_*...,
  z_offset)`` with ``add_source_class``, ``solve(dz, verbose, summary,
  progressbar)``, ``state``, ``res`` and a settable ``recomp_z_threshold``
  (``propagation.py:196-370, 813-890``).  ``__new__`` dispatches on
  ``prince_run.backend.linear_algebra_backend`` as the reference does
  (``propagation.py:398-404``); this package provides ``"hip"``.
* ``MultiRHSPropagationSolverETD2(..., K=)`` with per-column views
  ``solver[k].add_source_class`` / ``solver[k].res``
  (``propagation.py:2343-2562``).

The whole redshift march runs on the GPU.  The host prepares, once per
solve: the z grid and the refresh schedule (same float64 arithmetic as the
reference), per-step scalars (dl/dz, H(z) cm2sec, injection factors, source
evolution), per-window photon vectors and pair-production rates.  The device
folds the rate kernel for every window in one MFMA GEMM and then executes
two fused stage kernels per step (``libprince_hip.so``).
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from .. import _lib
from ..model.cosmology import H, hubble_steps
from ..model.partial_diff import DifferentialOperator
from ..model.units import PRINCE_UNITS
from .layout import (
    WindowLayout,
    build_block_layout,
    build_tasks,
    build_chunked_tasks,
    dense_split,
    fd_tables,
    fd_transposed_tables,
    transposed_layout,
)
from .result import UHECRPropagationResult
from .run_inputs import (
    em_cascade_grid,
    em_slices,
    fd_block,
    fold_kernel_fp64,
    pair_rate_per_energy,
    pair_species_scale,
)



def source_coefficients(src, zs):
    """Per-step coefficients ``norm * evolution(z)`` of one source over ``zs``."""
    f = getattr(src, "injection_coefficients", None)
    if f is not None:
        return f(zs)
    return np.array([src.norm * src.evolution(z) for z in zs], dtype=np.float64)

def z_grid_for(initial_z, final_z, dz):
    """Uniform z grid from initial_z to final_z (propagation.py:826-832)."""
    dz_step = -float(abs(dz))
    n_full = int(np.floor((final_z - initial_z) / dz_step))
    z_grid = initial_z + np.arange(n_full + 1) * dz_step
    if abs(z_grid[-1] - final_z) > 1e-12:
        z_grid = np.concatenate([z_grid, [final_z]])
    return z_grid


def refresh_schedule(z_grid, threshold):
    """(window index per step, refresh z per window) for the rate-cache gate.

    A step starting at z refreshes when no window exists yet or
    |z - z_cached| > threshold (propagation.py:1383-1387).
    """
    nsteps = len(z_grid) - 1
    win = np.empty(nsteps, dtype=np.int32)
    zs = []
    zc = None
    for k in range(nsteps):
        z0 = z_grid[k]
        if zc is None or abs(z0 - zc) > threshold:
            zc = z0
            zs.append(z0)
        win[k] = len(zs) - 1
    return win, np.asarray(zs, dtype=np.float64)


def _resolve_prince_run(args, kwargs):
    if "prince_run" in kwargs:
        return kwargs["prince_run"]
    if len(args) >= 3:
        return args[2]
    raise TypeError("ETD2 solver constructor requires `prince_run` (positional or keyword).")


def _torch():
    import torch

    return torch


class UHECRPropagationSolver:
    """Common solver state: sources, dl/dz, injection (propagation.py:195-370)."""

    def __init__(
        self,
        initial_z,
        final_z,
        prince_run,
        enable_adiabatic_losses=True,
        enable_pairprod_losses=True,
        enable_photohad_losses=True,
        enable_injection_jacobian=True,
        enable_partial_diff_jacobian=True,
        enable_decay=False,
        enable_decay_jacobian=True,
        z_offset=0.0,
    ):
        self.initial_z = initial_z + z_offset
        self.final_z = final_z + z_offset
        self.z_offset = z_offset
        self.prince_run = prince_run
        self.backend = prince_run.backend
        self.current_z_rates = None
        self.recomp_z_threshold = self.backend.update_rates_z_threshold
        self.spec_man = prince_run.spec_man
        self.egrid = prince_run.cr_grid.grid
        self.ebins = prince_run.cr_grid.bins
        self.widths = prince_run.cr_grid.widths
        self.enable_adiabatic_losses = enable_adiabatic_losses
        self.enable_pairprod_losses = enable_pairprod_losses
        self.enable_photohad_losses = enable_photohad_losses
        self.enable_injection_jacobian = enable_injection_jacobian
        self.enable_partial_diff_jacobian = enable_partial_diff_jacobian
        if enable_decay:
            raise NotImplementedError(
                "explicit decay (enable_decay=True) is not part of the HIP ETD2 path"
            )
        self.enable_decay = enable_decay
        self.enable_decay_jacobian = enable_decay_jacobian
        self.had_int_rates = prince_run.int_rates
        self.adia_loss_rates_grid = prince_run.adia_loss_rates_grid
        self.pair_loss_rates_grid = prince_run.pair_loss_rates_grid
        self.dim_states = prince_run.dim_states
        # co-evolved EM cascade (propagation.py:250-316): photon and e+- blocks
        # follow the nuclear states in ``state``
        em_grid = em_cascade_grid(prince_run)
        self.enable_em_cascade = em_grid is not None
        self.dim_em = 3 * em_grid.d if em_grid is not None else 0
        self._em_dz = 1e-2
        self.state = np.zeros(prince_run.dim_states + self.dim_em)
        self.result = None
        self.list_of_sources = []
        self.diff_operator = DifferentialOperator(prince_run.cr_grid, prince_run.spec_man.nspec)

    @property
    def known_species(self):
        return self.spec_man.known_species

    def _em_info(self):
        """{pdg: (grid, slice)} of the EM blocks of ``state`` (empty without EM)."""
        run = self.prince_run
        return {p: (run.em_grid.grid, sl) for p, sl in em_slices(run).items()}

    @property
    def res(self):
        if self.result is None:
            self.result = UHECRPropagationResult(self.state, self.egrid, self.spec_man,
                                                 em=self._em_info())
        return self.result

    def add_source_class(self, source_instance):
        self.list_of_sources.append(source_instance)

    def dldz(self, z):
        return -1.0 / ((1.0 + z) * H(z) * PRINCE_UNITS.cm2sec)

    def step_rates(self, z0):
        """``dldz(z)`` and ``H(z) cm2sec`` for every step start z of ``z0``,
        point by point: bit for bit the scalar ``dldz`` and ``H`` (evaluated
        on Python floats unless a subclass overrides ``dldz``)."""
        c2s = PRINCE_UNITS.cm2sec
        if type(self).dldz is UHECRPropagationSolver.dldz:
            hs = hubble_steps(z0)
            zl = np.asarray(z0, dtype=np.float64).tolist()
            dl = [-1.0 / ((1.0 + z) * h * c2s) for z, h in zip(zl, hs)]
            return np.array(dl, dtype=np.float64), np.array([h * c2s for h in hs], dtype=np.float64)
        return (np.array([self.dldz(z) for z in z0], dtype=np.float64),
                np.array([H(z) * c2s for z in z0], dtype=np.float64))

    def injection(self, dz, z):
        f = self.dldz(z) * dz * PRINCE_UNITS.cm2sec
        if len(self.list_of_sources) > 1:
            return f * np.sum([s.injection_rate(z) for s in self.list_of_sources], axis=0)
        return f * self.list_of_sources[0].injection_rate(z)


class UHECRPropagationSolverETD2(UHECRPropagationSolver):
    """ETD2 (Cox-Matthews exponential RK2) integrator facade.

    ``__new__`` picks the backend subclass from
    ``prince_run.backend.linear_algebra_backend``; ``"hip"`` gives
    :class:`ETD2SolverHIP`.  Other backends belong to the reference package.
    """

    def __new__(cls, *args, **kwargs):
        if cls is UHECRPropagationSolverETD2:
            prince_run = _resolve_prince_run(args, kwargs)
            backend = prince_run.backend.linear_algebra_backend.lower()
            if backend != "hip":
                raise NotImplementedError(
                    f"linear_algebra_backend={backend!r}: this package implements the 'hip' "
                    "backend only"
                )
            return object.__new__(ETD2SolverHIP)
        return object.__new__(cls)

    def solve(self, dz=1e-3, verbose=False, summary=False, progressbar=False):
        """Integrate from initial_z down to final_z with steps of |dz|."""
        start = time.time()
        z_grid = z_grid_for(self.initial_z, self.final_z, dz)
        self._em_dz = float(abs(dz))
        self.current_z_rates = None
        self._reset_for_solve()
        self._ensure_D_split()
        state = self._coerce_state(self._init_state())
        self._run_integration(state, z_grid)
        self.state = self._finalize_state(state)
        self.result = None
        end = time.time()
        self.last_solve_seconds = end - start
        if summary or verbose:
            print("ETD2 summary:")
            print(f"  steps: {len(z_grid) - 1}")
            print(f"  initial z: {self.initial_z} -> final z: {self.final_z}")
            print(f"  wall time: {end - start:.3f} s")


class _HipDeviceCache:
    """Per-run device residency of the rate kernel.

    When the run's rate object carries the kernel's Toeplitz factors
    (``int_rates.toeplitz``: per-channel response tables and one descriptor
    per kernel row, a quarter of the dense kernel's bytes for the A <= 56 run)
    only those are uploaded and every kernel row the solver needs is expanded
    on the device (``prh_toeplitz_rows``); otherwise the dense kernel
    ``_batch_matrix`` (the reference's object) is uploaded once and rows are
    gathered from it (``prh_gather_rows``).  Both give the same rows bit for
    bit.
    """

    def __init__(self, run, device):
        torch = _torch()
        self.device = device
        self.layouts = {}
        self.lib = _lib.load()
        ir = run.int_rates
        tz = getattr(ir, "toeplitz", None)
        self.tz = None
        self.B_full = None
        if tz is not None:
            desc = np.ascontiguousarray(tz.desc, dtype=np.int32)
            kind, tg, tl, i, k = (desc[:, c] for c in range(5))
            n1, n2 = tz.tab1d.shape[0], tz.tab2d.shape[0]
            ok = (np.all((kind == 0) | (kind == 1)) and np.all(tl < n1)
                  and np.all(np.where(kind == 0, tg < n1, (tg >= 0) & (tg < n2)))
                  and np.all((i >= 0) & (i < tz.d) & (k >= 0) & (k < tz.d)))
            if not ok or desc.shape[0] != ir.coupling_mat.nnz:
                raise ValueError("Toeplitz factors of the rate kernel are inconsistent")
            dv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
            self.tz_bufs = {"desc": dv(desc), "tab1d": dv(tz.tab1d), "tab2d": dv(tz.tab2d),
                            "fmo": dv(tz.fmo), "fdiff": dv(tz.fdiff)}
            b = self.tz_bufs
            self.tz = _lib.PrhToeplitz(
                d=tz.d, dph=tz.dph, n1=n1, n2=n2, nrows=desc.shape[0],
                desc=b["desc"].data_ptr(), tab1d=b["tab1d"].data_ptr(),
                tab2d=b["tab2d"].data_ptr(), fmo=b["fmo"].data_ptr(), fdiff=b["fdiff"].data_ptr())
            self.nrows, self.dph = desc.shape[0], tz.dph
            self.resident_bytes = int(sum(t.numel() * t.element_size() for t in b.values()))
        else:
            B = fold_kernel_fp64(ir)
            self.B_full = torch.from_numpy(B).to(device)
            self.nrows, self.dph = B.shape
            self.resident_bytes = B.nbytes

    def rows(self, sel_dev, nsel, out, stream):
        """Kernel rows sel (int64 device tensor, -1: zero row) into out [nsel][dph]."""
        import ctypes as _ct

        if self.tz is not None:
            _lib.check(self.lib.prh_toeplitz_rows(_ct.byref(self.tz), sel_dev.data_ptr(), nsel,
                                                  out.data_ptr(), stream), "kernel rows (Toeplitz)")
        else:
            _lib.check(self.lib.prh_gather_rows(self.B_full.data_ptr(), sel_dev.data_ptr(), nsel,
                                                self.dph, out.data_ptr(), stream), "kernel row gather")


class ETD2SolverHIP(UHECRPropagationSolverETD2):
    """ETD2 on one MI355X through the HIP C ABI (``libprince_hip.so``).

    Implements the reference's backend hook set (``_reset_for_solve``,
    ``_ensure_D_split``, ``_init_state``, ``_coerce_state``,
    ``_refresh_z_caches``, ``_ensure_apply_F``, ``_run_integration``,
    ``_finalize_state``, ``close``; ``propagation.py:430-459, 699-716``) with
    device-resident buffers: the host never touches the state between the
    first and last step.
    """

    #: steps per captured hipGraph chunk
    graph_chunk = 64

    def part_cost_for(self, K, pairs=False):
        """Cost units per row part (a diagonal block costs 3, a band block 2 +
        nslot); 48 with row-pair tasks (K = 512: 562 vs 585 ms per solve)."""
        tuned = getattr(self, "_tuned", {}).get(("part_cost", K))
        if tuned is not None:
            return tuned
        return 12 if K < 8 else (48 if pairs else 96)

    def wg_waves_for(self, K, with_adjoint=False, two_field=False):
        """Waves per stage workgroup: 8 from K = 64 on (more workgroups per CU:
        151 vs 165 us per stage at K = 512) and below K = 8 together with the
        dense-row pass (plain nuclear solve 115 ms); 16 in between, for
        adjoint plans (two-field forward at K = 64: 845 vs 1028 ms per
        checkpointed solve) and for two-field plans below K = 256 (K = 64 with
        2 right-hand sides per wave: 285 vs 311 ms per solve)."""
        tuned = getattr(self, "_tuned", {}).get(("wg_waves", K))
        if tuned is not None:
            return tuned
        if with_adjoint or (two_field and K < 256):
            return 16
        return 8 if (K >= 64 or K < 8) else 16

    def dense_rows_for(self, K):
        """How the dense rows (the few rows with a block from almost every
        species) are applied: "pass" = a mother-major kernel before the stage,
        "chunks" = cut into one-wave mother-chunk tasks inside the stage kernel
        plus a small epilogue launch, or None = ordinary split rows.  While
        the right-hand sides fill few workgroups the dense rows' long
        workgroups set the stage's critical path and "chunks" wins: plain
        nuclear solve 18.4 us per stage vs 22.1 ("pass") and 43.6 (split
        rows); K = 64: 57 vs 98 us.  At K = 512 the grid is full and split rows
        are faster (137 vs 165 us: the chunk tasks' partial sums and extra
        launch cost more than they save), so from K = 256 on the default is
        split rows; the autotuner tries both.  A tuned True means "pass"."""
        tuned = getattr(self, "_tuned", {}).get(("dense", K))
        if tuned is not None:
            return "pass" if tuned is True else (tuned or None)
        return "chunks" if K < 256 else None

    def dense_chunk_cost_for(self, K):
        """Cost units per mother chunk of a dense row in the "chunks" mode
        (24 below K = 8: 18.4 us per plain stage vs 22.6 with 12 and 19.5 with
        48)."""
        tuned = getattr(self, "_tuned", {}).get(("dense_chunk_cost", K))
        return tuned if tuned is not None else (24 if K < 8 else 96)

    def pair_rows_for(self, K):
        """Whether consecutive one-part diagonal rows run as row-pair tasks
        (one wave walks both rows' mothers merged and loads a shared mother
        panel once).  Only the one-field stages without dense-row modes take
        them; elsewhere the request is ignored.
        Default on from K = 256 (where the dense rows are ordinary split
        rows): 616 -> 562 ms per K = 512 solve, bitwise the same state."""
        tuned = getattr(self, "_tuned", {}).get(("pairs", K))
        return bool(tuned) if tuned is not None else K >= 256

    def adjoint_kt_for(self, K):
        """Right-hand sides per wave of the adjoint (transposed) stages."""
        tuned = getattr(self, "_tuned", {}).get(("kt_adjoint", K))
        return tuned if tuned is not None else self.kt_for(K)

    def adjoint_part_cost_for(self, K):
        """Part budget of the transposed operator's tasks: 24 from K = 8 on (twice
        the parts of the forward default; K = 64: backward 925 -> 850 ms, 12: 926)."""
        tuned = getattr(self, "_tuned", {}).get(("part_cost_adjoint", K))
        if tuned is not None:
            return tuned
        return self.part_cost_for(K) if K < 8 else 24

    def kt_for(self, K, with_adjoint=False, two_field=False):
        """Right-hand sides per wave: 8 from K = 256 on (K = 512: 109 / 118 vs
        118 / 123 us per stage with 4), else the largest of 4, 2, 1 not above K
        (K = 64: 38 / 39 vs 48 / 52 us with 8: twice the workgroups)."""
        tuned = getattr(self, "_tuned", {}).get(("kt", K))
        if tuned is not None:
            return tuned
        if (with_adjoint or two_field) and 2 <= K < 256:
            # two-field stages (16-wave workgroups) fill the chip better with 2:
            # K = 64 two-field solve 371 -> 285 ms (4 and 8-wave workgroups
            # before); an adjoint plan's forward 343 -> 286 ms, and the
            # backward's chunk recompute with it (with 4 in the transposed
            # stages: backward 850 -> 791 ms)
            return 2
        for kt in ((8, 4, 2, 1) if K >= 256 else (4, 2, 1)):
            if kt <= K:
                return kt
        return 1

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._lib = _lib.load()
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP ETD2 backend needs a ROCm GPU (torch.cuda is unavailable)")
        self.device = torch.device("cuda", int(getattr(self.backend, "device", 0)))
        self._K = 1
        self._plan = None
        self._D_tables = None
        self.timings = {}

    # ---------------------------------------------------------------- hooks
    def _reset_for_solve(self):
        self.close()

    def _ensure_D_split(self):
        if self._D_tables is None:
            self._fd_block = fd_block(self.diff_operator.operator, self.spec_man.nspec,
                                      self.prince_run.cr_grid.d)
            dcoef, ddiag = fd_tables(self._fd_block)
            self._D_tables = (dcoef, ddiag)

    def _init_state(self):
        n = self.dim_states + self.dim_em
        return np.zeros((n, self._K)) if self._K > 1 else np.zeros(n)

    def _coerce_state(self, state):
        """Host state (dim_states[, K]) to the device layout [nspec][K][d]."""
        torch = _torch()
        nspec, d = self.spec_man.nspec, self.prince_run.cr_grid.d
        K = self._K
        state = np.asarray(state, dtype=np.float64).reshape(-1, K)
        self._em_init = state[self.dim_states:] if self.dim_em else None
        host = state[: self.dim_states].reshape(nspec, d, K).transpose(0, 2, 1)
        return torch.from_numpy(np.ascontiguousarray(host)).to(self.device)

    def state_view(self, state):
        """Device state ([nspec][K][d]) as an [nspec, d, K] view."""
        return state.permute(0, 2, 1)

    def _finalize_state(self, state):
        nspec, d = self.spec_man.nspec, self.prince_run.cr_grid.d
        K = self._K
        host = self.state_view(state).cpu().numpy().reshape(nspec * d, K)
        if self.dim_em:
            host = np.concatenate([host, self.em_host_state()], axis=0)
        return np.ascontiguousarray(host[:, 0] if K == 1 else host)

    def em_host_state(self, plan=None):
        """EM blocks (e+, e-, photon) of the last solve as a host array [3 d_em, K]
        (the photons alternate between two buffers: after an odd number of
        steps they are in ``gam2``)."""
        plan = plan if plan is not None else self._plan
        eb = plan["em_bufs"]
        gam = (eb["gam2"] if plan.get("steps_run", plan["nsteps"]) % 2 else eb["gam"]).cpu().numpy()
        z = np.zeros_like(gam)
        return np.concatenate([z, z, gam], axis=0)

    def close(self):
        if self._plan is not None and self._plan.get("graph") is not None:
            self._plan["graph"].close()
        self._plan = None

    # ---------------------------------------------------------------- sources
    def _column_sources(self):
        """Sources per right-hand side (one column for the single-RHS solver)."""
        return [list(self.list_of_sources)]

    # ---------------------------------------------------------------- setup
    def _device_cache(self):
        run = self.prince_run
        cache = getattr(run, "_prh_device_cache", None)
        if cache is None or cache.device != self.device:
            cache = _HipDeviceCache(run, self.device)
            run._prh_device_cache = cache
        return cache

    def _layout_for(self, cache, phot0, with_transpose=False):
        """Window layout for the first-window photon vector (cached per run).

        The first window is folded on the device and read back once; the
        off-diagonal pattern is its non-zero set (the reference's
        ``eliminate_zeros`` at the first refresh).
        """
        torch = _torch()
        key = (phot0.tobytes(), bool(self.enable_photohad_losses), bool(with_transpose))
        lay = cache.layouts.get(key)
        if lay is not None:
            return lay
        stream = torch.cuda.current_stream(self.device).cuda_stream
        ph = torch.from_numpy(np.ascontiguousarray(phot0[None, :])).to(self.device)
        m0 = torch.empty(cache.nrows, dtype=torch.float64, device=self.device)
        full = cache.B_full
        if full is None:  # every row expanded once, for the first-window fold only
            full = torch.empty((cache.nrows, cache.dph), dtype=torch.float64, device=self.device)
            ident = torch.arange(cache.nrows, dtype=torch.int64, device=self.device)
            cache.rows(ident, cache.nrows, full, stream)
        _lib.check(
            self._lib.prh_fold_gemm(
                full.data_ptr(), cache.nrows, cache.dph, ph.data_ptr(), 1,
                m0.data_ptr(), cache.nrows, stream,
            ),
            "first-window fold",
        )
        m0_host = m0.cpu().numpy()
        del full
        ir = self.had_int_rates
        args = (ir.coupling_mat.indptr, ir.coupling_mat.indices, m0_host, self.spec_man.nspec,
                self.prince_run.cr_grid.d, bool(self.enable_photohad_losses))
        fwd = build_block_layout(*args)
        tr = transposed_layout(*args) if with_transpose else None
        wl = WindowLayout(fwd, tr)
        nsel = wl.sel.size
        sel_dev = torch.from_numpy(wl.sel).to(self.device)
        B_sel = torch.empty((nsel, cache.dph), dtype=torch.float64, device=self.device)
        cache.rows(sel_dev, nsel, B_sel, stream)
        lay = {"layout": wl, "B_sel": B_sel, "m0": m0_host}
        cache.layouts[key] = lay
        return lay

    # ------------------------------------------------------- photon-field split
    def set_field_scales(self, component, scales):
        """Scale one component of the run's combined photon field per right-hand side.

        Column k then sees n(E, z) = n_rest(E, z) + scales[k] * n_component(E, z):
        the rate matrix, its diagonal and the pair-production rate become
        ``M_rest + s_k M_component`` (two folds shared by every column).
        ``component`` indexes ``prince_run.photon_field.model_list``.
        """
        pf = self.prince_run.photon_field
        if not hasattr(pf, "model_list"):
            raise TypeError("set_field_scales needs a CombinedPhotonField")
        scales = np.asarray(scales, dtype=np.float64).reshape(-1)
        if scales.size != self._K:
            raise ValueError(f"need {self._K} scales, got {scales.size}")
        self._field_split = (int(component), scales.copy())

    def _split_fields(self):
        fs = getattr(self, "_field_split", None)
        if fs is None:
            return None
        idx, scales = fs
        pf = self.prince_run.photon_field
        comp = pf.model_list[idx]
        rest = [m for i, m in enumerate(pf.model_list) if i != idx]
        return rest, comp, scales

    def _refresh_z_caches(self, zs):
        """Inputs of every refresh window: photon vectors and pair rates.

        Returns (phot, rpair, phot_e, rpair_e); the last two are None unless a
        per-column field scale is set.  For a CMB [+ synthetic EBL] field they
        are evaluated on the device (``DeviceFieldTables``, tensors in HBM);
        other fields are evaluated by their host models.
        """
        from .field_tables import DeviceFieldTables, device_field

        parts = device_field(self.prince_run.photon_field)
        split = self._split_fields()
        if parts is not None and (split is None or (parts[1] is not None
                                                    and split[1] is parts[1])):
            ft = getattr(self.prince_run, "_prh_field_tables", None)
            if ft is None or ft.device != self.device:
                ft = DeviceFieldTables(self.prince_run, self.device)
                self.prince_run._prh_field_tables = ft
            if split is None:
                phot, rpair = ft.tables(zs)
                self._rpair_raw = rpair.clone()
                if not self.enable_pairprod_losses:
                    rpair.zero_()
                return phot, rpair, None, None
            self._rpair_raw = None
            phot, rpair = ft.tables(zs, ebl=False)
            phot_e, rpair_e = ft.tables(zs, cmb=False)
            if not self.enable_pairprod_losses:
                rpair.zero_()
                rpair_e.zero_()
            return phot, rpair, phot_e, rpair_e
        self._rpair_raw = None
        ir = self.had_int_rates
        pair = self.pair_loss_rates_grid
        d = self.prince_run.cr_grid.d
        split = self._split_fields()
        if split is None:
            phot = np.stack([ir.photon_vector(z) for z in zs]).astype(np.float64)
            if self.enable_pairprod_losses:
                rpair = np.stack([pair_rate_per_energy(pair, z) for z in zs])
            else:
                rpair = np.zeros((len(zs), d))
            return phot, rpair, None, None
        rest, comp, _ = split

        class _Sum:
            def get_photon_density(self_, E, z):
                E = np.atleast_1d(np.asarray(E, dtype=np.float64))
                return np.sum([m.get_photon_density(E, z) for m in rest], axis=0)

        base = _Sum()
        phot = np.stack([ir.photon_vector(z, pfield=base) for z in zs]).astype(np.float64)
        phot_e = np.stack([ir.photon_vector(z, pfield=comp) for z in zs]).astype(np.float64)
        if self.enable_pairprod_losses:
            rpair = np.stack([pair_rate_per_energy(pair, z, pfield=base) for z in zs])
            rpair_e = np.stack([pair_rate_per_energy(pair, z, pfield=comp) for z in zs])
        else:
            rpair = np.zeros((len(zs), d))
            rpair_e = np.zeros((len(zs), d))
        return phot, rpair, phot_e, rpair_e

    def _schedule_arrays(self, z_grid, col_sources):
        nsteps = len(z_grid) - 1
        z0 = z_grid[:-1]
        h = np.diff(z_grid)
        dldz, hub = self.step_rates(z0)
        if self.enable_adiabatic_losses and self.enable_partial_diff_jacobian:
            hadia = hub
        else:
            hadia = np.zeros(nsteps)
        finj = dldz * 1.0 * PRINCE_UNITS.cm2sec
        J = max(1, max(len(c) for c in col_sources))
        K = len(col_sources)
        coef = np.zeros((nsteps, J, K))
        for kcol, srcs in enumerate(col_sources):
            for j, src in enumerate(srcs):
                coef[:, j, kcol] = source_coefficients(src, z0)
        return h, dldz, hadia, finj, coef, J

    def _source_tables(self, col_sources, J):
        nspec, d = self.spec_man.nspec, self.prince_run.cr_grid.d
        K = len(col_sources)
        used = np.zeros(nspec, dtype=bool)
        for srcs in col_sources:
            for src in srcs:
                g = np.asarray(src.injection_grid).reshape(nspec, d)
                used |= np.any(g != 0.0, axis=1)
        slots = np.flatnonzero(used)
        src_slot = np.full(nspec, -1, dtype=np.int32)
        src_slot[slots] = np.arange(slots.size, dtype=np.int32)
        grid = np.zeros((max(1, slots.size), J, K, d))
        for kcol, srcs in enumerate(col_sources):
            for j, src in enumerate(srcs):
                g = np.asarray(src.injection_grid).reshape(nspec, d)
                grid[: slots.size, j, kcol, :] = g[slots]
        return src_slot, grid, slots

    def _ensure_apply_F(self, z_grid, with_adjoint=False):
        """Build the device plan of a solve: operator(s), schedule and buffers."""
        torch = _torch()
        t0 = time.perf_counter()
        dev = self.device
        K = self._K
        nspec, d = self.spec_man.nspec, self.prince_run.cr_grid.d
        if d > 64:
            raise NotImplementedError("the HIP stage kernels support at most 64 energy bins")
        cache = self._device_cache()
        win, zs = refresh_schedule(z_grid, self.recomp_z_threshold)
        phot, rpair, phot_e, rpair_e = self._refresh_z_caches(zs)
        two = phot_e is not None
        phot0 = phot[0] + phot_e[0] if two else phot[0]
        if not isinstance(phot0, np.ndarray):
            phot0 = phot0.cpu().numpy()
        lay = self._layout_for(cache, phot0, with_transpose=with_adjoint)
        wl = lay["layout"]
        nwin = len(zs)
        stride = wl.stride
        stream = torch.cuda.current_stream(dev).cuda_stream
        t1 = time.perf_counter()
        if with_adjoint and self.enable_em_cascade:
            raise NotImplementedError("the adjoint sweep does not cover the EM cascade")
        split_fields = self._split_fields() is not None
        wgw = self.wg_waves_for(K, with_adjoint, split_fields)
        dense = chunked = None
        mode = self.dense_rows_for(K)
        if mode and not with_adjoint and not split_fields and self.kt_for(K) <= 8:
            if mode == "pass":
                dense = dense_split(wl.blocks, wl.fwd.brow_ptr, nspec)
            elif mode == "chunks":
                chunked = build_chunked_tasks(
                    wl.fwd, self.part_cost_for(K), wgw, chunk_cost=self.dense_chunk_cost_for(K),
                    place=getattr(self, "_tuned", {}).get(("dense_place", K), "mother"))
            else:
                raise ValueError(f"unknown dense-row mode {mode!r}")
        if dense is not None:
            from dataclasses import replace as _dc_replace

            rest = _dc_replace(wl.fwd, blocks=dense[3], brow_ptr=dense[4])
            tasks = build_tasks(rest, self.part_cost_for(K), wgw)
        elif chunked is not None:
            tasks = chunked[0]
        else:
            pairs = (self.pair_rows_for(K) and not with_adjoint and not split_fields
                     and self.kt_for(K) <= 8)
            tasks = build_tasks(wl.fwd, self.part_cost_for(K, pairs), wgw, pair_rows=pairs)
        kt = self.kt_for(K, with_adjoint, split_fields)
        col_sources = self._column_sources()
        has_src = any(len(c) for c in col_sources) and self.enable_injection_jacobian
        h, dldz, hadia, finj, coef, J = self._schedule_arrays(z_grid, col_sources)
        src_slot, src_grid, slots = self._source_tables(col_sources, J)
        dcoef, ddiag = self._D_tables
        pscale = pair_species_scale(self.pair_loss_rates_grid, self.spec_man)

        def dv(a, dtype=None):
            if isinstance(a, torch.Tensor):
                return a.contiguous()
            a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
            return torch.from_numpy(a).to(dev)

        flags = 0
        if self.enable_partial_diff_jacobian:
            flags |= _lib.PRH_F_PARTIAL_DIFF
        if has_src:
            flags |= _lib.PRH_F_INJECTION
        if np.any(np.asarray(tasks)[:, 5] & _lib.PRH_TASK_PAIR):
            flags |= _lib.PRH_F_ROW_PAIRS
        blocks = wl.blocks if dense is None else dense[3]
        blocks = blocks if blocks.size else np.zeros((1, 4))
        sshape = (nspec, K, d)
        bufs = {
            "tasks": dv(tasks, np.int32),
            "blocks": dv(blocks, np.int32),
            "mvals": torch.empty((nwin, stride), dtype=torch.float64, device=dev),
            "phot": dv(phot, np.float64),
            "B_sel": lay["B_sel"],
            "egrid": dv(self.prince_run.cr_grid.grid, np.float64),
            "dcoef": dv(dcoef, np.float64),
            "ddiag": dv(ddiag, np.float64),
            "pscale": dv(pscale, np.float64),
            "rpair": dv(rpair, np.float64),
            "src_slot": dv(src_slot, np.int32),
            "src_grid": dv(src_grid, np.float64),
            "h": dv(h, np.float64),
            "dldz": dv(dldz, np.float64),
            "hadia": dv(hadia, np.float64),
            "finj": dv(finj, np.float64),
            "win": dv(win, np.int32),
            "src_coef": dv(coef, np.float64),
            "a": torch.zeros(sshape, dtype=torch.float64, device=dev),
            "fphi": torch.zeros(sshape, dtype=torch.float64, device=dev),
            "step_ctl": torch.zeros(2, dtype=torch.int32, device=dev),
        }
        DENSE_CHUNK = getattr(self, "_tuned", {}).get(("dense_chunk", K), 4 if K < 8 else 16)
        ndchunk = -(-nspec // DENSE_CHUNK)
        if dense is not None:
            bufs["dense_row"] = dv(dense[1], np.int32)
            bufs["dense_tab"] = dv(dense[2], np.int32)
            bufs["dpart"] = torch.empty((ndchunk, len(dense[0]), K, d), dtype=torch.float64, device=dev)
        if chunked is not None:
            _, fin_tasks, drows, row_of, ndchunk = chunked
            bufs["dense_row"] = dv(row_of, np.int32)
            bufs["fin_tasks"] = dv(fin_tasks, np.int32)
            bufs["dpart"] = torch.empty((ndchunk, len(drows), K, d), dtype=torch.float64, device=dev)
        if two:
            bufs["mvals_e"] = torch.empty((nwin, stride), dtype=torch.float64, device=dev)
            bufs["phot_e"] = dv(phot_e, np.float64)
            bufs["rpair_e"] = dv(rpair_e, np.float64)
            bufs["escale"] = dv(self._field_split[1], np.float64)

        def make_op(tk, blk, dcf, wgw=16, dense_op=False, kt=kt):
            return _lib.PrhOperator(
                nspec=nspec, d=d, K=K, ntasks=tk.shape[0],
                tasks=tk.data_ptr() if tk.numel() else None,
                blocks=blk.data_ptr(), mvals_stride=stride, diag_off=wl.diag_off,
                mvals=bufs["mvals"].data_ptr(), egrid=bufs["egrid"].data_ptr(),
                dcoef=dcf.data_ptr(), ddiag=bufs["ddiag"].data_ptr(),
                pscale=bufs["pscale"].data_ptr(), rpair=bufs["rpair"].data_ptr(),
                src_slot=bufs["src_slot"].data_ptr(), nsrc=J, flags=flags,
                src_grid=bufs["src_grid"].data_ptr(), kt=kt, pad_=0,
                mvals_e=bufs["mvals_e"].data_ptr() if two else None,
                rpair_e=bufs["rpair_e"].data_ptr() if two else None,
                escale=bufs["escale"].data_ptr() if two else None,
                wg_waves=wgw,
                ndense=(len(dense[0]) if dense is not None else len(chunked[2]) if chunked is not None
                        else 0) if dense_op else 0,
                dchunk=DENSE_CHUNK, ndchunk=ndchunk,
                dense_row=bufs["dense_row"].data_ptr() if ("dense_row" in bufs and dense_op) else None,
                dense_tab=bufs["dense_tab"].data_ptr() if (dense is not None and dense_op) else None,
                nfin=int(chunked[1].shape[0]) if (chunked is not None and dense_op) else 0, pad3_=0,
                fin_tasks=bufs["fin_tasks"].data_ptr() if (chunked is not None and dense_op) else None,
            )

        op = make_op(bufs["tasks"], bufs["blocks"], bufs["dcoef"], wgw=wgw, dense_op=True)
        opT = None
        if with_adjoint:
            tasks_t = build_tasks(wl.tr, self.adjoint_part_cost_for(K), _lib.WG_WAVES)
            bufs["tasks_t"] = dv(tasks_t, np.int32)
            bufs["blocks_t"] = dv(wl.blocks_t if wl.blocks_t.size else np.zeros((1, 4)), np.int32)
            bufs["dcoef_t"] = dv(fd_transposed_tables(self._fd_block), np.float64)
            opT = make_op(bufs["tasks_t"], bufs["blocks_t"], bufs["dcoef_t"],
                          kt=self.adjoint_kt_for(K))
        sch = _lib.PrhSchedule(
            nsteps=len(h), pad_=0, h=bufs["h"].data_ptr(), dldz=bufs["dldz"].data_ptr(),
            hadia=bufs["hadia"].data_ptr(), finj=bufs["finj"].data_ptr(),
            win=bufs["win"].data_ptr(), src_coef=bufs["src_coef"].data_ptr(),
        )
        x = torch.zeros(sshape, dtype=torch.float64, device=dev)
        em_struct, em_bufs = (None, None)
        if self.enable_em_cascade:
            em_struct, em_bufs = self._em_setup(zs, K, dv)
            de = self.prince_run.em_grid.d
            em_bufs["qstep"] = torch.empty((len(h), de, de), dtype=torch.float64, device=dev)
            em_struct.qstep = em_bufs["qstep"].data_ptr()
            _lib.check(self._lib.prh_em_step_transfer(
                ctypes.byref(op), ctypes.byref(sch), ctypes.byref(em_struct),
                em_bufs["qstep"].data_ptr(), stream), "EM step transfers")
        st = _lib.PrhState(
            x=x.data_ptr(), a=bufs["a"].data_ptr(), fphi=bufs["fphi"].data_ptr(),
            step_ctl=bufs["step_ctl"].data_ptr(), rec_x=None, rec_a=None, rec_fx=None,
            rec_fa=None, rec_first=0, rec_count=0,
            em=ctypes.addressof(em_struct) if em_struct is not None else None,
            dpart=bufs["dpart"].data_ptr() if "dpart" in bufs else None,
        )
        self._plan = {
            "op": op, "opT": opT, "sch": sch, "st": st, "x": x, "bufs": bufs,
            "nsteps": len(h), "layout": wl, "nwin": nwin, "zs": zs, "tasks": tasks, "kt": kt,
            "stream": stream, "stride": stride, "dph": cache.dph, "K": K,
            "fold_rows": stride, "graph": None, "two": two, "z_grid": z_grid,
            "src_slots": slots, "J": J, "win": win,
            "dense_rows": (list(dense[0]) if dense is not None
                           else list(chunked[2]) if chunked is not None else []),
            "dense_mode": "pass" if dense is not None else "chunks" if chunked is not None else None,
            "em": em_struct, "em_bufs": em_bufs,
        }
        if getattr(self.backend, "use_hip_graphs", True) and len(h) >= self.graph_chunk:
            self._plan["graph"] = _lib.StepGraph(op, sch, st, self.graph_chunk)
        self.refresh_zs = zs
        self.timings["setup_host_s"] = t1 - t0
        return self._plan

    def _em_setup(self, zs, K, dv):
        """Device buffers of the EM sector: per-window cascade transfers T(z) and
        Bethe-Heitler shapes at the refresh redshifts (propagation.py:461-618),
        built on the device (``DeviceCascade``), the per-window pair weights and
        the photon state."""
        from ..model.cascade import cascade_kernels
        from ..model.partial_diff import single_block

        from .cascade_device import DeviceCascade

        torch = _torch()
        run = self.prince_run
        dev = self.device
        if self.enable_decay:
            raise NotImplementedError("explicit decay with the EM cascade")
        if self._split_fields() is not None:
            raise NotImplementedError("per-column photon-field scales with the EM cascade need "
                                      "the uneven batch (per-column cascade transfers)")
        eg = run.em_grid
        if eg.d > _lib.EM_MAX:
            raise NotImplementedError(f"the EM kernels take at most {_lib.EM_MAX} bins")
        ck = cascade_kernels(eg.grid, run.cr_grid.grid)
        field = run.photon_field
        dc = getattr(run, "_prh_device_cascade", None)
        if dc is None or dc.device != self.device:
            dc = DeviceCascade(ck, self.device)
            run._prh_device_cascade = dc
        Tg, Te, Rbh, iters = dc.build(zs, field, self._em_dz)
        E_p = run.cr_grid.grid
        dE = np.gradient(E_p)
        pair = self.pair_loss_rates_grid
        # Bethe-Heitler weights dl_step (dE/E) rate(z) per window, with the pair
        # rate whether or not the nuclear pair loss is on (the reference's hook)
        dl = np.array([PRINCE_UNITS.c * self._em_dz / ((1.0 + z) * H(z)) for z in zs])
        raw = getattr(self, "_rpair_raw", None)
        if raw is not None:
            bhw = (torch.from_numpy(dl).to(dev)[:, None]
                   * torch.from_numpy(dE / E_p).to(dev)[None, :]) * raw
        else:
            bhw = np.stack([dl[w] * (dE / E_p) * pair_rate_per_energy(pair, z)
                            for w, z in enumerate(zs)])
        pscale = pair_species_scale(pair, self.spec_man)
        cbh = np.zeros(self.spec_man.nspec)
        for sref in self.spec_man.species_refs:
            if sref.is_nucleus and abs(sref.charge) >= 1:
                cbh[sref.princeidx] = float(sref.A) * pscale[sref.princeidx]
        dcoef, ddiag = fd_tables(single_block(eg))
        gam0 = np.zeros((eg.d, K))
        init = getattr(self, "_em_init", None)
        if init is not None and init.size:
            gam0 = np.asarray(init[2 * eg.d : 3 * eg.d], dtype=np.float64).reshape(eg.d, K)
        b = {
            "egrid": dv(eg.grid, np.float64), "dcoef": dv(dcoef, np.float64),
            "ddiag": dv(ddiag, np.float64), "cbh": dv(cbh, np.float64),
            "bhw": dv(bhw, np.float64), "rbh": Rbh, "tg": Tg, "te": Te, "iters": iters,
            "gam": dv(gam0, np.float64), "gam0": dv(gam0, np.float64),
            "gam2": torch.zeros((eg.d, K), dtype=torch.float64, device=dev),
            "wbh": torch.zeros((run.cr_grid.d, K), dtype=torch.float64, device=dev),
        }
        d = run.cr_grid.d
        b["tstack"] = torch.empty((len(zs), eg.d, eg.d + d), dtype=torch.float64, device=dev)
        _lib.check(self._lib.prh_em_transfer_stack(
            Tg.data_ptr(), Te.data_ptr(), Rbh.data_ptr(), len(zs), eg.d, d,
            b["tstack"].data_ptr(), torch.cuda.current_stream(dev).cuda_stream), "EM transfer stack")
        em = _lib.PrhEm(
            d_em=eg.d, nwin=len(zs), egrid=b["egrid"].data_ptr(), dcoef=b["dcoef"].data_ptr(),
            ddiag=b["ddiag"].data_ptr(), cbh=b["cbh"].data_ptr(), bhw=b["bhw"].data_ptr(),
            rbh=b["rbh"].data_ptr(), tg=b["tg"].data_ptr(), te=b["te"].data_ptr(),
            gam=b["gam"].data_ptr(), wbh=b["wbh"].data_ptr(), gam2=b["gam2"].data_ptr(),
            tstack=b["tstack"].data_ptr(),
        )
        self.em_iterations = iters.cpu().numpy().tolist()
        return em, b

    # ---------------------------------------------------------------- device run
    def prepare(self, dz=1e-3, with_adjoint=False):
        """Host-side setup of a solve; every input ends up resident in HBM."""
        self._em_dz = float(abs(dz))
        z_grid = z_grid_for(self.initial_z, self.final_z, dz)
        self._reset_for_solve()
        self._ensure_D_split()
        return self._ensure_apply_F(z_grid, with_adjoint=with_adjoint)

    def fold_windows(self, plan):
        """Rate fold of every refresh window (one MFMA GEMM per field) into mvals."""
        b = plan["bufs"]
        pairs = [("phot", "mvals")] + ([("phot_e", "mvals_e")] if plan["two"] else [])
        for ph, mv in pairs:
            _lib.check(
                self._lib.prh_fold_gemm(
                    b["B_sel"].data_ptr(), plan["fold_rows"], plan["dph"], b[ph].data_ptr(),
                    plan["nwin"], b[mv].data_ptr(), plan["stride"], plan["stream"],
                ),
                "window fold",
            )
            if not self.enable_photohad_losses:
                b[mv].zero_()

    def execute(self, plan=None, k_begin=0, k_end=None):
        """Device part of a solve: zero state, fold, ETD2 steps (asynchronous)."""
        plan = plan if plan is not None else self._plan
        if k_begin == 0:
            plan["x"].zero_()
            self._reset_em_state(plan)
            self.fold_windows(plan)
        k_end = plan["nsteps"] if k_end is None else k_end
        self._launch_steps(plan, k_begin, k_end)
        return plan["x"]

    @staticmethod
    def _reset_em_state(plan):
        """Photon state back to the solve's initial EM state (step 0 reads ``gam``)."""
        eb = plan.get("em_bufs")
        if eb is not None:
            eb["gam"].copy_(eb["gam0"])

    # ---------------------------------------------------------------- autotuner
    def tuning_candidates(self, K):
        """Stage configurations the autotuner times for K right-hand sides:
        right-hand sides per wave (kt), part budget (block cost per wave part,
        i.e. rows per wave), waves per workgroup, the dense-row mode and, with
        dense rows as split rows, row-pair tasks on or off."""
        kts = [kt for kt in (8, 4, 2, 1) if kt <= K][:2]
        out = []
        for kt in kts:
            for pc in (48, 96) if K >= 8 else (12, 24):
                for wgw in (8, 16):
                    for dense in (("chunks", False) if K >= 8 else (True, "chunks", False)):
                        for pairs in ((True, False) if dense is False else (False,)):
                            out.append({("kt", K): kt, ("part_cost", K): pc,
                                        ("wg_waves", K): wgw, ("dense", K): dense,
                                        ("pairs", K): pairs})
        return out

    def autotune(self, dz=1e-3, steps=30, candidates=None, repeats=2):
        """Time each candidate stage configuration on this operator for ``steps``
        ETD2 steps (HIP events on the launch stream, after one untimed pass) and
        keep the fastest in ``self._tuned``.  Returns the timing table
        [(seconds per step, config)], fastest first.  Results do not depend on
        the choice beyond summation order (every candidate is parity-tested)."""
        torch = _torch()
        K = self._K
        cands = candidates if candidates is not None else self.tuning_candidates(K)
        table = []
        for cfg in cands:
            self._tuned = dict(cfg)
            try:
                plan = self.prepare(dz)
            except (ValueError, NotImplementedError):
                continue
            n = min(steps, plan["nsteps"])
            plan["x"].zero_()
            self.fold_windows(plan)
            self._launch_steps(plan, 0, n, use_graph=False)
            best = None
            for _ in range(repeats):
                plan["x"].zero_()
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                ev0.record()
                self._launch_steps(plan, 0, n, use_graph=False)
                ev1.record()
                torch.cuda.synchronize(self.device)
                t = ev0.elapsed_time(ev1) * 1e-3 / n
                best = t if best is None else min(best, t)
            table.append((best, cfg))
        self.close()
        if not table:
            raise RuntimeError("no stage configuration could be prepared")
        table.sort(key=lambda r: r[0])
        self._tuned = dict(table[0][1])
        self.tuning_table = table
        return table

    def _launch_steps(self, plan, k_begin, k_end, use_graph=True):
        plan["steps_run"] = k_end  # the EM photons' buffer after the last step
        if use_graph and plan["graph"] is not None:
            plan["graph"].run(k_begin, k_end, plan["stream"])
            return
        _lib.check(
            self._lib.prh_etd2_steps(
                ctypes.byref(plan["op"]), ctypes.byref(plan["sch"]), ctypes.byref(plan["st"]),
                k_begin, k_end, plan["stream"],
            ),
            "ETD2 steps",
        )

    # ---------------------------------------------------------------- adjoint
    def forward_checkpointed(self, plan, chunk=None):
        """Forward solve keeping the state at every ``chunk``-th step (default: the
        graph chunk, 64) for the adjoint sweep; returns the checkpoint tensor."""
        torch = _torch()
        chunk = chunk or self.graph_chunk
        nsteps = plan["nsteps"]
        nchk = (nsteps + chunk - 1) // chunk
        ck = plan.get("ckpt")
        if ck is None or ck.shape[0] != nchk:
            ck = torch.empty((nchk,) + tuple(plan["x"].shape), dtype=torch.float64,
                             device=self.device)
            plan["ckpt"] = ck
        plan["ckpt_chunk"] = chunk
        plan["x"].zero_()
        self._reset_em_state(plan)
        self.fold_windows(plan)
        for c in range(nchk):
            ck[c].copy_(plan["x"])
            self._launch_steps(plan, c * chunk, min((c + 1) * chunk, nsteps),
                               use_graph=(chunk == self.graph_chunk))
        return ck

    def _adjoint_buffers(self, plan):
        torch = _torch()
        ab = plan.get("adj_bufs")
        if ab is not None:
            return ab
        dev = self.device
        shape = tuple(plan["x"].shape)
        C = plan["ckpt_chunk"]
        J, K, d = plan["J"], plan["K"], self.prince_run.cr_grid.d
        nslot = max(1, len(plan["src_slots"]))
        z0 = plan["z_grid"][:-1]
        ab = {
            "lam": torch.zeros(shape, dtype=torch.float64, device=dev),
            "lam_a": torch.zeros(shape, dtype=torch.float64, device=dev),
            "u": torch.zeros(shape, dtype=torch.float64, device=dev),
            "v": torch.zeros(shape, dtype=torch.float64, device=dev),
            "rec": torch.zeros((4, C) + shape, dtype=torch.float64, device=dev),
            "gsrc0": torch.zeros((nslot, J, K, d), dtype=torch.float64, device=dev),
            "gsrc1": torch.zeros((nslot, J, K, d), dtype=torch.float64, device=dev),
            "gdot": torch.zeros(shape, dtype=torch.float64, device=dev),
            "lnz1": torch.from_numpy(np.log1p(z0)).to(dev),
        }
        rec = ab["rec"]
        plan["st_rec"] = _lib.PrhState(
            x=plan["x"].data_ptr(), a=plan["bufs"]["a"].data_ptr(),
            fphi=plan["bufs"]["fphi"].data_ptr(), step_ctl=plan["bufs"]["step_ctl"].data_ptr(),
            rec_x=rec[0].data_ptr(), rec_a=rec[1].data_ptr(), rec_fx=rec[2].data_ptr(),
            rec_fa=rec[3].data_ptr(), rec_first=0, rec_count=C,
        )
        plan["adj"] = _lib.PrhAdjoint(
            lam=ab["lam"].data_ptr(), lam_a=ab["lam_a"].data_ptr(), u=ab["u"].data_ptr(),
            v=ab["v"].data_ptr(), rec_x=rec[0].data_ptr(), rec_a=rec[1].data_ptr(),
            rec_fx=rec[2].data_ptr(), rec_fa=rec[3].data_ptr(), rec_first=0, rec_count=C,
            gsrc0=ab["gsrc0"].data_ptr(), gsrc1=ab["gsrc1"].data_ptr(),
            lnz1=ab["lnz1"].data_ptr(),
            gdot=ab["gdot"].data_ptr() if plan["two"] else None,
        )
        if getattr(self.backend, "use_hip_graphs", True) and C == self.graph_chunk:
            plan["rec_graph"] = _lib.StepGraph(plan["op"], plan["sch"], plan["st_rec"], C)
        else:
            plan["rec_graph"] = None
        plan["adj_bufs"] = ab
        return ab

    def adjoint_sweep(self, plan, lam_final):
        """Reverse ETD2 sweep from dJ/dx(final) (device tensor, state layout).

        Recomputes each checkpoint chunk forward (recording x, a, F(x), F(a)),
        then runs the adjoint steps of the chunk backwards.  Returns the
        accumulator tensors: ``gsrc0`` / ``gsrc1`` ([slot][J][K][d]: dJ/db
        summed with dl/dz cm2sec * norm * evolution and with an extra
        ln(1+z)) and ``gdot`` ([nspec][K][d], per-element terms of dJ/ds for
        the scaled photon-field component), and ``lam`` = dJ/dx(initial).
        """
        if plan.get("opT") is None:
            raise RuntimeError("plan built without the transposed operator (prepare(with_adjoint=True))")
        if plan.get("ckpt") is None:
            raise RuntimeError("run forward_checkpointed(plan) first")
        ab = self._adjoint_buffers(plan)
        for key in ("gsrc0", "gsrc1", "gdot", "u", "v", "lam_a"):
            ab[key].zero_()
        ab["lam"].copy_(lam_final)
        nsteps = plan["nsteps"]
        C = plan["ckpt_chunk"]
        ck = plan["ckpt"]
        adj = plan["adj"]
        opT = plan["opT"]
        sch = plan["sch"]
        stream = plan["stream"]
        lib = self._lib
        _lib.check(lib.prh_adjoint_seed(ctypes.byref(opT), ctypes.byref(sch), ctypes.byref(adj),
                                        nsteps - 1, stream), "adjoint seed")
        for c in range((nsteps + C - 1) // C - 1, -1, -1):
            k0, k1 = c * C, min((c + 1) * C, nsteps)
            plan["x"].copy_(ck[c])
            if plan["rec_graph"] is not None:
                plan["rec_graph"].run(k0, k1, stream)
            else:
                st = plan["st_rec"]
                st.rec_first = k0
                _lib.check(lib.prh_etd2_steps(ctypes.byref(plan["op"]), ctypes.byref(sch),
                                              ctypes.byref(st), k0, k1, stream), "recompute")
            adj.rec_first = k0
            _lib.check(lib.prh_adjoint_steps(ctypes.byref(opT), ctypes.byref(sch),
                                             ctypes.byref(adj), k0, k1, stream), "adjoint steps")
        return ab

    def _run_integration(self, state, z_grid, step_hook=None):
        torch = _torch()
        plan = self._ensure_apply_F(z_grid)
        plan["x"].copy_(state)
        self._reset_em_state(plan)
        self.fold_windows(plan)
        self._launch_steps(plan, 0, plan["nsteps"])
        state.copy_(plan["x"])
        torch.cuda.synchronize(self.device)
        return state


# =====================================================================
# Multi-RHS: K independent solutions sharing the operator
# =====================================================================


class _MultiRHSView:
    """Column k of a multi-RHS solver (propagation.py:2343-2372)."""

    def __init__(self, parent, k):
        self.parent = parent
        self.k = k
        self.list_of_sources = []

    def add_source_class(self, source_instance):
        self.list_of_sources.append(source_instance)

    @property
    def state(self):
        s = self.parent.state
        return s[:, self.k] if s.ndim == 2 else s

    @property
    def res(self):
        return UHECRPropagationResult(self.state, self.parent.egrid, self.parent.spec_man,
                                      em=self.parent._em_info())


class MultiRHSPropagationSolverETD2(UHECRPropagationSolverETD2):
    """K right-hand sides sharing one operator and one fold (propagation.py:2375-2562)."""

    def __new__(cls, *args, **kwargs):
        if cls is MultiRHSPropagationSolverETD2:
            prince_run = _resolve_prince_run(args, kwargs)
            backend = prince_run.backend.linear_algebra_backend.lower()
            if backend != "hip":
                raise NotImplementedError(
                    f"linear_algebra_backend={backend!r}: this package implements the 'hip' "
                    "backend only"
                )
            return object.__new__(_MultiRHSETD2SolverHIP)
        return object.__new__(cls)

    def __init__(self, *args, K, **kwargs):
        if K < 1:
            raise ValueError(f"K must be >= 1, got {K!r}")
        self._K_req = int(K)
        super().__init__(*args, **kwargs)
        self._K = int(K)
        self._views = [_MultiRHSView(self, k) for k in range(self._K)]
        self.state = np.zeros((self.dim_states, self._K))

    def __len__(self):
        return self._K

    def __getitem__(self, k):
        return self._views[k]

    @property
    def K(self):
        return self._K

    def add_source_class(self, source_instance):
        raise TypeError(
            "MultiRHSPropagationSolverETD2 expects per-RHS sources via "
            "solver[k].add_source_class(...)."
        )

    @property
    def list_of_sources(self):
        for v in getattr(self, "_views", []):
            if v.list_of_sources:
                return v.list_of_sources
        return []

    @list_of_sources.setter
    def list_of_sources(self, value):
        if value:
            raise TypeError("Use solver[k].add_source_class(...) to attach sources.")

    @property
    def res(self):
        raise TypeError("MultiRHSPropagationSolverETD2.res is per-view: use solver[k].res")

    def injection(self, dz, z):
        f = self.dldz(z) * dz * PRINCE_UNITS.cm2sec
        out = np.zeros((self.dim_states, self._K))
        for k, view in enumerate(self._views):
            srcs = view.list_of_sources
            if not srcs:
                continue
            col = (
                np.sum([s.injection_rate(z) for s in srcs], axis=0)
                if len(srcs) > 1
                else srcs[0].injection_rate(z)
            )
            out[:, k] = f * col
        return out


class _MultiRHSETD2SolverHIP(MultiRHSPropagationSolverETD2, ETD2SolverHIP):
    """Multi-RHS on the GPU: the stage kernels run one wave per (row part, column)."""

    def _column_sources(self):
        return [list(v.list_of_sources) for v in self._views]
