"""Uneven batch: K independent ETD2 solves with their own initial redshifts and
photon-field scales, advanced together on one GPU.

Column k is exactly the reference solve ``UHECRPropagationSolverETD2(z_max[k],
final_z, run)`` (``propagation.py:813-890``) on the field ``CMB + s_k * EBL``:
its own z grid (``propagation.py:826-832``), its own refresh windows
(``propagation.py:1383-1387``) and therefore its own coupling values.  The
device keeps one window of coupling values per column ([K][window layout]);
at every step the columns that cross a window boundary are re-folded
(``prh_fold_rows``: one MFMA GEMM over just those columns), then both stages run
for every column that has steps left (``prh_uneven_step``); finished columns
keep their final state.  Columns are ordered by decreasing step count, so the
workgroups of finished columns leave at once.

Per-window photon vectors and pair-loss rates of every column come from
``prh_field_tables`` (the CMB Planck spectrum evaluated on the device plus the
z-separable EBL tabulated at z = 0).

With the EM cascade (``prince_run.enable_em_cascade``) every column also
carries photons on the EM grid: the transfer T_gamma, T_electron and the
Bethe-Heitler shapes of every column window are built on the device before
the march (``DeviceCascade``: the column's field CMB + s_k EBL at the
window's redshift), and after the nuclear stages of a step each live column
runs its own EM step (``prh_uneven_step`` with ``prh_state.em``): the
reference's step hook (``propagation.py:867-878``) per column.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from .. import _lib
from ..model import config as mconfig
from ..model.cosmology import hubble_steps
from ..model.photonfields import PLANCK_PREF, CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL
from ..model.units import PRINCE_UNITS
from .layout import build_tasks
from .run_inputs import pair_species_scale
from .propagation import (
    MultiRHSPropagationSolverETD2,
    _MultiRHSETD2SolverHIP,
    _torch,
    refresh_schedule,
    source_coefficients,
    z_grid_for,
)


def _field_parts(field):
    """(CMB component, EBL component) of a CMB + SyntheticEBL field."""
    if not isinstance(field, CombinedPhotonField):
        raise NotImplementedError("the uneven batch needs a CombinedPhotonField(CMB, SyntheticEBL)")
    cmb = [m for m in field.model_list if isinstance(m, CMBPhotonSpectrum)]
    ebl = [m for m in field.model_list if isinstance(m, SyntheticEBL)]
    if len(cmb) != 1 or len(ebl) != 1 or len(field.model_list) != 2:
        raise NotImplementedError("the uneven batch needs exactly one CMB and one SyntheticEBL")
    return cmb[0], ebl[0]


def column_field(field, scale):
    """The field a column with EBL scale ``scale`` sees (host form, for checks)."""
    _, ebl = _field_parts(field)
    return CombinedPhotonField(
        [CMBPhotonSpectrum(), SyntheticEBL(amplitudes=ebl.amplitudes, scale=ebl.scale * scale)]
    )


class UnevenBatchSolverETD2(_MultiRHSETD2SolverHIP):
    """K columns with initial redshifts ``z_max[k]`` and EBL scales ``ebl_scale[k]``.

    Sources per column via ``solver[k].add_source_class(src)``; ``solve(dz)``
    leaves the final states in ``state`` ([dim_states, K], column order as
    given).
    """

    #: right-hand sides per wave of the uneven stage kernel
    kt_uneven = 2
    #: value slots per column: 2 lets the windows that start within a group of
    #: steps be folded together before the group (one pass over the fold
    #: kernel per group instead of per step); 1 folds each window in place at
    #: its first step
    value_slots = 2

    def __new__(cls, *args, **kwargs):
        return object.__new__(cls)

    def __init__(self, z_max, prince_run, final_z=0.0, ebl_scale=None, **kwargs):
        z_max = np.asarray(z_max, dtype=np.float64).reshape(-1)
        K = z_max.size
        if np.any(z_max <= final_z):
            raise ValueError("every z_max must lie above final_z")
        MultiRHSPropagationSolverETD2.__init__(self, float(z_max.max()), final_z, prince_run, K=K,
                                               **kwargs)
        self.z_max = z_max
        self.ebl_scale = (np.ones(K) if ebl_scale is None
                          else np.asarray(ebl_scale, dtype=np.float64).reshape(-1))
        if self.ebl_scale.size != K:
            raise ValueError("need one EBL scale per column")
        _field_parts(prince_run.photon_field)

    # ---------------------------------------------------------------- schedule
    def _column_schedules(self, dz):
        """Per-column z grids and windows, ordered by decreasing step count."""
        cols = []
        for k in range(self._K):
            zg = z_grid_for(float(self.z_max[k]), self.final_z, dz)
            win, zs = refresh_schedule(zg, self.recomp_z_threshold)
            cols.append((zg, np.asarray(win), np.asarray(zs)))
        order = np.argsort([-(len(c[0]) - 1) for c in cols], kind="stable")
        return cols, order

    @staticmethod
    def _refold_group(starts, smax=64):
        """Largest S (a power of two up to smax) such that no column starts two
        windows within one aligned group of S steps: then the windows starting
        in a group can be folded before it, each into the slot its
        predecessor's predecessor has left."""
        T = starts.shape[0]
        S = smax
        while S > 1:
            pad = (-T) % S
            grp = np.concatenate([starts, np.zeros((pad, starts.shape[1]), dtype=bool)])
            if grp.reshape(-1, S, starts.shape[1]).sum(axis=1).max() <= 1:
                return S
            S //= 2
        return 1

    def _device_tables(self, zw, sw, dev):
        """Photon vectors [nwin][dph] and pair rates [nwin][d] of all column windows."""
        from .field_tables import DeviceFieldTables

        return DeviceFieldTables(self.prince_run, dev).tables(zw, scales=sw)

    def prepare(self, dz=1e-3, with_adjoint=False):
        if with_adjoint:
            raise NotImplementedError("the uneven batch has no adjoint")
        torch = _torch()
        t0 = time.perf_counter()
        self._reset_for_solve()
        self._ensure_D_split()
        dev = self.device
        K = self._K
        run = self.prince_run
        nspec, d = self.spec_man.nspec, run.cr_grid.d
        cols, order = self._column_schedules(dz)
        nsteps = np.array([len(cols[k][0]) - 1 for k in order])
        T = int(nsteps.max())
        # per-step tables in device column order
        h = np.zeros((T, K))
        dldz = np.zeros((T, K))
        hadia = np.zeros((T, K))
        finj = np.zeros((T, K))
        gwin = np.full((T, K), -1, dtype=np.int32)
        col_sources = self._column_sources()
        col_sources = [col_sources[k] for k in order]
        J = max(1, max(len(c) for c in col_sources))
        coef = np.zeros((T, J, K))
        zw, sw = [], []
        ref_in, ref_out, ref_off = [], [], [0]
        off = 0
        firsts = []
        for c, k in enumerate(order):
            zg, win, zs = cols[k]
            n = len(zg) - 1
            z0 = zg[:-1]
            h[:n, c] = np.diff(zg)
            # scalar evaluation per step, as the plain path (numpy's vector pow and
            # sqrt round differently from the scalar ones)
            dl, hub = self.step_rates(z0)
            dldz[:n, c] = dl
            if self.enable_adiabatic_losses and self.enable_partial_diff_jacobian:
                hadia[:n, c] = hub
            finj[:n, c] = dl * 1.0 * PRINCE_UNITS.cm2sec
            gwin[:n, c] = off + win
            for j, src in enumerate(col_sources[c]):
                coef[:n, j, c] = source_coefficients(src, z0)
            zw.extend(zs)
            sw.extend([self.ebl_scale[k]] * len(zs))
            firsts.append(off)
            off += len(zs)
        live = np.count_nonzero(gwin >= 0, axis=1).astype(np.int32)
        if not all(np.all(gwin[t, :live[t]] >= 0) for t in range(T)):
            raise AssertionError("live columns must lead (columns ordered by decreasing steps)")
        # re-fold lists: the windows starting at step t, or with two value
        # slots all windows starting in the group of S steps from t
        starts = np.zeros((T, K), dtype=bool)
        starts[0] = gwin[0] >= 0
        starts[1:] = (gwin[1:] >= 0) & (gwin[1:] != gwin[:-1])
        nslot = 2 if self.value_slots == 2 else 1
        S = self._refold_group(starts) if nslot == 2 else 1
        for t in range(T):
            if t % S == 0:
                tt, cc = np.nonzero(starts[t:t + S])
                w = gwin[t + tt, cc]
                ref_in.extend(w.tolist())
                ref_out.extend(((w & 1) * K + cc if nslot == 2 else cc).tolist())
            ref_off.append(len(ref_in))
        zw = np.asarray(zw)
        sw = np.asarray(sw)
        phot, rpair = self._device_tables(zw, sw, dev)
        # the EM sector's Bethe-Heitler weights use the pair rate whether or not
        # the nuclear pair loss is switched on (as the reference's hook)
        rpair_em = rpair.clone() if self.enable_em_cascade else None
        if not self.enable_pairprod_losses:
            rpair.zero_()
        # operator pattern from the first window of the first column; every
        # column's first window must have the same non-zero photon bins
        ph_host = phot.cpu().numpy()
        nz0 = ph_host[firsts[0]] != 0
        if any(not np.array_equal(ph_host[f] != 0, nz0) for f in firsts):
            raise NotImplementedError("columns whose first windows differ in their non-zero photon bins")
        cache = self._device_cache()
        lay = self._layout_for(cache, ph_host[firsts[0]])
        wl = lay["layout"]
        stride = wl.stride
        tasks = build_tasks(wl.fwd, self.part_cost_for(K), _lib.WG_WAVES)
        kt = min(self.kt_uneven, 4)
        while kt > K:
            kt //= 2
        src_slot, src_grid, slots = self._source_tables(col_sources, J)
        dcoef, ddiag = self._D_tables
        pscale = pair_species_scale(self.pair_loss_rates_grid, self.spec_man)
        has_src = any(len(c) for c in col_sources) and self.enable_injection_jacobian

        def dv(a, dtype=None):
            a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
            return torch.from_numpy(a).to(dev)

        flags = 0
        if self.enable_partial_diff_jacobian:
            flags |= _lib.PRH_F_PARTIAL_DIFF
        if has_src:
            flags |= _lib.PRH_F_INJECTION
        bufs = {
            "tasks": dv(tasks, np.int32),
            "blocks": dv(wl.blocks if wl.blocks.size else np.zeros((1, 4)), np.int32),
            "B_sel": lay["B_sel"], "phot": phot, "rpair": rpair,
            "mrhs": torch.zeros((nslot * K, stride), dtype=torch.float64, device=dev),
            "egrid": dv(run.cr_grid.grid, np.float64), "dcoef": dv(dcoef, np.float64),
            "ddiag": dv(ddiag, np.float64), "pscale": dv(pscale, np.float64),
            "src_slot": dv(src_slot, np.int32), "src_grid": dv(src_grid, np.float64),
            "h": dv(h, np.float64), "dldz": dv(dldz, np.float64), "hadia": dv(hadia, np.float64),
            "finj": dv(finj, np.float64), "win": dv(gwin, np.int32), "src_coef": dv(coef, np.float64),
            "ref_in": dv(np.asarray(ref_in + [0]), np.int32),
            "ref_out": dv(np.asarray(ref_out + [0]), np.int32),
            "a": torch.zeros((nspec, K, d), dtype=torch.float64, device=dev),
            "fphi": torch.zeros((nspec, K, d), dtype=torch.float64, device=dev),
            "step_ctl": torch.zeros(2, dtype=torch.int32, device=dev),
        }
        x = torch.zeros((nspec, K, d), dtype=torch.float64, device=dev)
        op = _lib.PrhOperator(
            nspec=nspec, d=d, K=K, ntasks=tasks.shape[0], tasks=bufs["tasks"].data_ptr(),
            blocks=bufs["blocks"].data_ptr(), mvals_stride=stride, diag_off=wl.diag_off,
            mvals=bufs["mrhs"].data_ptr(), egrid=bufs["egrid"].data_ptr(),
            dcoef=bufs["dcoef"].data_ptr(), ddiag=bufs["ddiag"].data_ptr(),
            pscale=bufs["pscale"].data_ptr(), rpair=bufs["rpair"].data_ptr(),
            src_slot=bufs["src_slot"].data_ptr(), nsrc=J, flags=flags,
            src_grid=bufs["src_grid"].data_ptr(), kt=kt, pad_=0,
        )
        un = _lib.PrhUneven(
            nsteps=T, nwin=int(zw.size), h=bufs["h"].data_ptr(), dldz=bufs["dldz"].data_ptr(),
            hadia=bufs["hadia"].data_ptr(), finj=bufs["finj"].data_ptr(),
            win=bufs["win"].data_ptr(), src_coef=bufs["src_coef"].data_ptr(),
            rpair=bufs["rpair"].data_ptr(), mrhs=bufs["mrhs"].data_ptr(), nslot=nslot,
            live=live.ctypes.data,
        )
        st = _lib.PrhState(x=x.data_ptr(), a=bufs["a"].data_ptr(), fphi=bufs["fphi"].data_ptr(),
                           step_ctl=bufs["step_ctl"].data_ptr())
        if self.enable_em_cascade:
            em_struct, em_bufs = self._em_tables(zw, sw, rpair_em, order, dz, dev, dv)
            st.em = ctypes.addressof(em_struct)
        else:
            em_struct = em_bufs = None
        self._plan = {
            "op": op, "un": un, "st": st, "x": x, "bufs": bufs, "nsteps": T, "order": order,
            "ref_off": np.asarray(ref_off), "stride": stride, "dph": cache.dph, "K": K,
            "stream": torch.cuda.current_stream(dev).cuda_stream, "layout": wl, "kt": kt,
            "nwin": int(zw.size), "col_steps": nsteps, "graph": None, "em_bufs": em_bufs,
            "live": live, "refold_group": S, "nslot": nslot,
            "em": em_struct, "two": False,
        }
        self.timings["setup_host_s"] = time.perf_counter() - t0
        return self._plan

    def _em_tables(self, zw, sw, rpair, order, dz, dev, dv):
        """Device EM tables of every column window (in device column order):
        T_gamma, T_electron, R_BH, Bethe-Heitler weights; photon state."""
        from ..model.cascade import cascade_kernels
        from ..model.partial_diff import single_block
        from .cascade_device import DeviceCascade
        from .layout import fd_tables
        from .run_inputs import pair_species_scale

        torch = _torch()
        run = self.prince_run
        eg = run.em_grid
        if eg.d > _lib.EM_MAX:
            raise NotImplementedError(f"the EM kernels take at most {_lib.EM_MAX} bins")
        ck = cascade_kernels(eg.grid, run.cr_grid.grid)
        cache = getattr(run, "_prh_device_cascade", None)
        if cache is None or cache.device != dev:
            cache = DeviceCascade(ck, dev)
            run._prh_device_cascade = cache
        t0 = time.perf_counter()
        tg, te, rbh, iters = cache.build(zw, run.photon_field, dz, scales=sw)
        E_p = run.cr_grid.grid
        dl = PRINCE_UNITS.c * abs(dz) / ((1.0 + zw) * np.array(hubble_steps(zw)))
        dEoE = np.gradient(E_p) / E_p
        bhw = (torch.from_numpy(dl).to(dev)[:, None] * torch.from_numpy(dEoE).to(dev)[None, :]) * rpair
        pscale = pair_species_scale(self.pair_loss_rates_grid, self.spec_man)
        cbh = np.zeros(self.spec_man.nspec)
        for sref in self.spec_man.species_refs:
            if sref.is_nucleus and abs(sref.charge) >= 1:
                cbh[sref.princeidx] = float(sref.A) * pscale[sref.princeidx]
        dcoef, ddiag = fd_tables(single_block(eg))
        b = {
            "egrid": dv(eg.grid, np.float64), "dcoef": dv(dcoef, np.float64),
            "ddiag": dv(ddiag, np.float64), "cbh": dv(cbh, np.float64),
            "bhw": bhw.contiguous(), "rbh": rbh, "tg": tg, "te": te, "iters": iters,
            "gam": torch.zeros((eg.d, self._K), dtype=torch.float64, device=dev),
        }
        em = _lib.PrhEm(
            d_em=eg.d, nwin=int(zw.size), egrid=b["egrid"].data_ptr(), dcoef=b["dcoef"].data_ptr(),
            ddiag=b["ddiag"].data_ptr(), cbh=b["cbh"].data_ptr(), bhw=b["bhw"].data_ptr(),
            rbh=b["rbh"].data_ptr(), tg=b["tg"].data_ptr(), te=b["te"].data_ptr(),
            gam=b["gam"].data_ptr(), wbh=None, gam2=None,
        )
        torch.cuda.synchronize(dev)
        self.timings["em_build_s"] = time.perf_counter() - t0
        return em, b

    def _chunk_graphs(self, plan):
        """Capture the solve as 64-step chunk graphs (re-folds and stages of every
        step baked in); built once per plan."""
        if plan.get("graphs") is not None:
            return plan["graphs"]
        b = plan["bufs"]
        ro = np.ascontiguousarray(plan["ref_off"], dtype=np.int32)
        plan["ref_off_i32"] = ro
        rf = _lib.PrhRefold(
            kern=b["B_sel"].data_ptr(), nrows=plan["stride"], dph=plan["dph"],
            enabled=1 if self.enable_photohad_losses else 0, phot=b["phot"].data_ptr(),
            in_rows=b["ref_in"].data_ptr(), out_rows=b["ref_out"].data_ptr(),
            ref_off=ro.ctypes.data,
        )
        plan["refold"] = rf
        T, C = plan["nsteps"], self.graph_chunk
        plan["graphs"] = [
            _lib.UnevenChunkGraph(plan["op"], plan["un"], plan["st"], rf, c0, min(c0 + C, T))
            for c0 in range(0, T, C)
        ]
        return plan["graphs"]

    def close(self):
        if self._plan is not None:
            for g in self._plan.get("graphs") or []:
                g.close()
            self._plan["graphs"] = None
        super().close()

    def execute(self, plan=None, k_begin=0, k_end=None):
        """Device part of the batch: zero state, then per step the re-folds and both
        stages (replayed as 64-step chunk graphs when the whole range is run)."""
        plan = plan if plan is not None else self._plan
        b = plan["bufs"]
        lib = self._lib
        stream = plan["stream"]
        k_end = plan["nsteps"] if k_end is None else k_end
        if k_begin == 0:
            plan["x"].zero_()
            if plan.get("em_bufs") is not None:
                plan["em_bufs"]["gam"].zero_()
        C = self.graph_chunk
        if (getattr(self.backend, "use_hip_graphs", True) and k_begin % C == 0
                and (k_end % C == 0 or k_end == plan["nsteps"])):
            graphs = self._chunk_graphs(plan)
            for g in graphs[k_begin // C: -(-k_end // C)]:
                g.launch(stream)
            return plan["x"]
        ro = plan["ref_off"]
        isz = ctypes.sizeof(ctypes.c_int32)
        for t in range(k_begin, k_end):
            n = int(ro[t + 1] - ro[t])
            if n and self.enable_photohad_losses:
                _lib.check(lib.prh_fold_rows(
                    b["B_sel"].data_ptr(), plan["stride"], plan["dph"], b["phot"].data_ptr(),
                    b["ref_in"].data_ptr() + int(ro[t]) * isz, b["ref_out"].data_ptr() + int(ro[t]) * isz,
                    n, b["mrhs"].data_ptr(), plan["stride"], stream), "column re-fold")
            _lib.check(lib.prh_uneven_step(ctypes.byref(plan["op"]), ctypes.byref(plan["un"]),
                                           ctypes.byref(plan["st"]), t, stream), "uneven step")
        return plan["x"]

    def solve(self, dz=1e-3, verbose=False, summary=False, progressbar=False):
        torch = _torch()
        start = time.time()
        plan = self.prepare(dz)
        self.execute(plan)
        torch.cuda.synchronize(self.device)
        nspec, d, K = self.spec_man.nspec, self.prince_run.cr_grid.d, self._K
        host = plan["x"].cpu().numpy().transpose(0, 2, 1).reshape(nspec * d, K)
        if plan.get("em_bufs") is not None:
            gam = plan["em_bufs"]["gam"].cpu().numpy()
            z = np.zeros_like(gam)
            host = np.concatenate([host, z, z, gam], axis=0)  # e+, e-, photon blocks
        state = np.empty_like(host)
        state[:, plan["order"]] = host
        self.state = state
        self.result = None
        self.last_solve_seconds = time.time() - start
        if summary or verbose:
            print(f"uneven batch: {K} columns, {plan['nsteps']} steps, {plan['nwin']} windows, "
                  f"{self.last_solve_seconds:.3f} s")
        return state
