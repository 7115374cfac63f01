"""Differentiable propagation: adjoint ETD2 gradients of a flux/composition loss.

No counterpart exists in the reference, which has no adjoint.  Problem: K source
instances on the nuclear run, each with 8 parameters

    theta_k = (gamma, log10 R_max, l_1..l_4, m, s_EBL)

composition fractions softmax(l_1, l_2, l_3, l_4, 0) over {H, He, N, Si, Fe},
an ``AugerFitSource`` with evolution (1+z)^m, and the EBL component of the
photon field scaled by s_EBL (differentiated through the rate fold).  Loss per
instance:

    sum_{E > 10^9.6 GeV} ((log10 J - log10 J_t)/0.05)^2 + ((<lnA> - <lnA>_t)/0.1)^2

with J the all-particle flux and <lnA> the mean log mass on the common total-
energy grid of ``UHECRPropagationResult.get_solution_group`` / ``get_lnA``
(reference ``solvers/propagation.py:103-186``, log-log interpolation with
non-positive knots dropped).

The forward solve keeps a state checkpoint every 64 steps; the backward sweep
recomputes each chunk and runs the hand-written adjoint kernels
(``prh_adjoint_step``).  Gradients w.r.t. the source parameters come from the
accumulated dJ/db contracted with the analytic derivatives of the injection
grids; the EBL-scale gradient is accumulated inside the adjoint kernels.
"""

from __future__ import annotations

import numpy as np

from . import workloads
from .model.cr_sources import AugerFitSource
from .model.units import EnergyGrid
from .model import config
from .solvers.propagation import MultiRHSPropagationSolverETD2

NPARAM = 8
EBL_COMPONENT = 1  # index of the EBL in CombinedPhotonField([CMB, EBL])


def _torch():
    import torch

    return torch


def fractions(logits4):
    z = np.concatenate([np.asarray(logits4, dtype=np.float64), [0.0]])
    z = z - z.max()
    e = np.exp(z)
    return e / e.sum()


def make_source(run, theta):
    gamma, lgr = float(theta[0]), float(theta[1])
    f = fractions(theta[2:6])
    params = {pid: (gamma, 10.0**lgr, float(fk)) for pid, fk in zip(workloads.COMPOSITION, f)}
    return AugerFitSource(run, params=params, m=("simple", float(theta[6])))


def draw_params(K, seed=0):
    """Seeded parameter draws in the batched-workload ranges (EBL scale U[0.6, 1.4])."""
    rng = np.random.default_rng(seed)
    th = np.zeros((K, NPARAM))
    th[:, 0] = rng.uniform(-1.5, 2.5, K)
    th[:, 1] = rng.uniform(8.5, 10.5, K)
    th[:, 2:6] = rng.normal(0.0, 1.0, (K, 4))
    th[:, 6] = rng.uniform(-6.0, 6.0, K)
    th[:, 7] = rng.uniform(0.6, 1.4, K)
    return th


def source_derivatives(run, theta):
    """Per species of the composition: g_p and dg_p/d(gamma, log10 R_max) on the grid."""
    gamma, lgr = float(theta[0]), float(theta[1])
    f = fractions(theta[2:6])
    E = run.cr_grid.grid
    out = []
    for pid, fp in zip(workloads.COMPOSITION, f):
        sp = run.spec_man.pdgid2sref[pid]
        A, Z = float(sp.A), sp.Z
        ek = A * E
        emax = Z * 10.0**lgr
        cut = np.where(ek < emax, 1.0, np.exp(1 - ek / emax))
        g = fp * A * (ek / 1e9) ** (-gamma) * cut
        dg_dgamma = -np.log(ek / 1e9) * g
        dg_dlgr = np.where(ek < emax, 0.0, g * (ek / emax) * np.log(10.0))
        out.append((sp.princeidx, g, dg_dgamma, dg_dlgr, fp))
    return out


class Observables:
    """All-particle flux and <ln A> on the common total-energy grid (torch, differentiable).

    Mirrors ``_collect_interpolated_spectra``: species spectra in total energy
    (epow = 0), log-log interpolated onto the common grid, non-positive knots
    dropped, zero outside a species' positive range.  The bracketing knots are
    fixed from the state the observables are first evaluated on.
    """

    def __init__(self, run, emin_log=9.6):
        sm = run.spec_man
        self.A = np.array([s.A for s in sm.species_refs], dtype=np.float64)
        lo, hi, bpd = config.cosmic_ray_grid
        self.egrid = EnergyGrid(lo, np.log10(self.A.max()) + hi, bpd * 4).grid
        self.cr = run.cr_grid.grid
        self.emask = self.egrid > 10**emin_log
        self.ec = self.egrid[self.emask]
        step = np.log(self.cr[1] / self.cr[0])
        # fractional knot coordinate of every common-grid point for every species
        self.t = (np.log(self.ec)[None, :] - np.log(self.A)[:, None] - np.log(self.cr[0])) / step

    def interp_plan(self, state_np):
        """state_np: [nspec, d, K] -> knot indices and weights ([nspec, nc, K])."""
        nspec, d, K = state_np.shape
        pos = state_np > 0.0
        idx = np.arange(d)[None, :, None]
        prev = np.where(pos, idx, -1)
        prev = np.maximum.accumulate(prev, axis=1)
        nxt = np.where(pos, idx, d)
        nxt = np.flip(np.minimum.accumulate(np.flip(nxt, axis=1), axis=1), axis=1)
        t = self.t[:, :, None] * np.ones((1, 1, K))
        i = np.clip(np.floor(t).astype(np.int64), 0, d - 2)
        inside = (t >= 0) & (t <= d - 1)
        si = np.arange(nspec)[:, None, None]
        ki = np.arange(K)[None, None, :]
        L = prev[si, i, ki]
        R = nxt[si, np.minimum(i + 1, d - 1), ki]
        exact = np.isclose(t, np.round(t), rtol=0, atol=1e-9) & pos[si, np.clip(np.round(t).astype(np.int64), 0, d - 1), ki]
        valid = inside & (L >= 0) & (R < d)
        L = np.where(valid, L, 0)
        R = np.where(valid, R, 1)
        w = np.where(valid, (t - L) / np.maximum(R - L, 1), 0.0)
        tr = np.clip(np.round(t).astype(np.int64), 0, d - 1)
        L = np.where(exact, tr, L)
        R = np.where(exact, tr, R)
        w = np.where(exact, 0.0, w)
        valid = valid | (exact & inside)
        return L, R, w, valid

    def evaluate(self, state_t, plan):
        """state_t: torch [nspec, d, K] -> (log10 J [nc, K], <lnA> [nc, K], J [nc, K])."""
        torch = _torch()
        L, R, w, valid = (torch.as_tensor(a, device=state_t.device) for a in plan)
        nspec, d, K = state_t.shape
        si = torch.arange(nspec, device=state_t.device)[:, None, None]
        ki = torch.arange(K, device=state_t.device)[None, None, :]
        A = torch.as_tensor(self.A, device=state_t.device)
        nL = state_t[si, L, ki]
        nR = state_t[si, R, ki]
        safeL = torch.where(valid, nL, torch.ones_like(nL))
        safeR = torch.where(valid, nR, torch.ones_like(nR))
        logv = (1 - w) * torch.log(safeL) + w * torch.log(safeR) - torch.log(A)[:, None, None]
        spec = torch.where(valid, torch.exp(logv), torch.zeros_like(logv))
        J = spec.sum(dim=0)
        lnA = (torch.log(A)[:, None, None] * spec).sum(dim=0) / torch.where(J > 0, J, torch.ones_like(J))
        return torch.log10(torch.where(J > 0, J, torch.ones_like(J))), lnA, J


def loss_terms(log10J, lnA, J, tgt_log10J, tgt_lnA, mask):
    torch = _torch()
    a = torch.where(mask, (log10J - tgt_log10J) / 0.05, torch.zeros_like(log10J))
    b = torch.where(mask, (lnA - tgt_lnA) / 0.1, torch.zeros_like(lnA))
    return (a * a + b * b).sum(dim=0)


class AdjointProblem:
    """Batched forward + adjoint for K instances on the nuclear run (z_max -> 0)."""

    def __init__(self, K, z_max=3.0, dz=1e-3, run=None, target_seed=1234):
        self.run = run if run is not None else workloads.nuclear_run()
        self.K = int(K)
        self.z_max = float(z_max)
        self.dz = float(dz)
        self.obs = Observables(self.run)
        self.target_seed = target_seed
        self._target = None

    # --------------------------------------------------------------- forward
    def _solver(self, theta):
        s = MultiRHSPropagationSolverETD2(self.z_max, 0.0, self.run, K=self.K)
        for k in range(self.K):
            s[k].add_source_class(make_source(self.run, theta[k]))
        s.set_field_scales(EBL_COMPONENT, theta[:, 7])
        return s

    def forward(self, theta, with_adjoint=True):
        """Solve for parameters theta [K, 8]; returns (solver, plan, state [nspec, d, K] torch)."""
        theta = np.asarray(theta, dtype=np.float64)
        s = self._solver(theta)
        plan = s.prepare(self.dz, with_adjoint=with_adjoint)
        if with_adjoint:
            s.forward_checkpointed(plan)
        else:
            s.execute(plan)
        x = plan["x"]  # [nspec, K, d]
        return s, plan, x.permute(0, 2, 1)

    def target(self):
        if self._target is None:
            torch = _torch()
            th = draw_params(self.K, seed=self.target_seed)
            _, _, st = self.forward(th, with_adjoint=False)
            stn = st.detach().cpu().numpy()
            ip = self.obs.interp_plan(stn)
            l10, lna, J = self.obs.evaluate(st.detach(), ip)
            mask = J > 0
            self._target = (l10.detach(), lna.detach(), mask, th)
            torch.cuda.synchronize()
        return self._target

    def loss_of_state(self, state_t, interp=None, mask=None):
        """Per-instance losses (torch) of a state [nspec, d, K].

        ``interp`` (bracketing knots) and ``mask`` (bins entering the loss)
        default to the ones of ``state_t``; pass those of a base state to keep
        the loss smooth across nearby states (finite differences).
        """
        tl10, tlna, tmask, _ = self.target()
        if interp is None:
            interp = self.obs.interp_plan(state_t.detach().cpu().numpy())
        l10, lna, J = self.obs.evaluate(state_t, interp)
        if mask is None:
            mask = tmask & (J > 0)
        return loss_terms(l10, lna, J, tl10, tlna, mask), (interp, mask)

    def loss(self, theta):
        """Per-instance losses (numpy) from a forward solve only."""
        _, _, st = self.forward(theta, with_adjoint=False)
        losses, _ = self.loss_of_state(st)
        return losses.detach().cpu().numpy()

    # --------------------------------------------------------------- gradient
    def loss_and_grad(self, theta):
        """(losses [K], dloss/dtheta [K, 8], timings) via the adjoint sweep."""
        import time

        torch = _torch()
        theta = np.asarray(theta, dtype=np.float64)
        t0 = time.perf_counter()
        solver, plan, st = self.forward(theta, with_adjoint=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stv = st.detach().clone().requires_grad_(True)
        losses, self.last_loss_plan = self.loss_of_state(stv)
        losses.sum().backward()
        lam_final = stv.grad.permute(0, 2, 1).contiguous()  # back to [nspec, K, d]
        ab = solver.adjoint_sweep(plan, lam_final)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grad = self.contract(solver, plan, ab, theta)
        t3 = time.perf_counter()
        timings = {"forward_s": t1 - t0, "backward_s": t2 - t1, "contract_s": t3 - t2}
        return losses.detach().cpu().numpy(), grad, timings

    def contract(self, solver, plan, ab, theta):
        """Parameter gradients from the adjoint accumulators."""
        K = self.K
        d = self.run.cr_grid.d
        slots = list(plan["src_slots"])
        g0 = ab["gsrc0"].cpu().numpy()  # [slot, J, K, d]
        g1 = ab["gsrc1"].cpu().numpy()
        gdot = ab["gdot"].cpu().numpy()  # [nspec, K, d]
        grad = np.zeros((K, NPARAM))
        for k in range(K):
            derivs = source_derivatives(self.run, theta[k])
            f = fractions(theta[k, 2:6])
            dot_g = np.zeros(len(derivs))
            for p, (sidx, g, dgg, dgr, fp) in enumerate(derivs):
                sl = slots.index(sidx)
                G0 = g0[sl, 0, k]
                G1 = g1[sl, 0, k]
                grad[k, 0] += G0 @ dgg
                grad[k, 1] += G0 @ dgr
                grad[k, 6] += G1 @ g
                dot_g[p] = G0 @ g
            # d g_p / d l_j = g_p (delta_pj - f_j) for the four free logits
            for j in range(4):
                grad[k, 2 + j] = dot_g[j] - f[j] * dot_g.sum()
            grad[k, 7] = gdot[:, k, :].sum()
        return grad


class ETD2Propagate:
    """``torch.autograd.Function`` factory: theta [K, 8] -> final state [nspec, d, K].

    ``apply(theta, problem)`` runs the checkpointed forward; ``backward``
    runs the adjoint sweep from the incoming state gradient and contracts the
    accumulators into d/dtheta.
    """

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        torch = _torch()

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta, problem):
                th = theta.detach().cpu().numpy()
                solver, plan, st = problem.forward(th, with_adjoint=True)
                ctx.problem, ctx.solver, ctx.plan, ctx.theta = problem, solver, plan, th
                ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
                return st.detach().clone()

            @staticmethod
            def backward(ctx, grad_state):
                lam_final = grad_state.permute(0, 2, 1).contiguous()
                ab = ctx.solver.adjoint_sweep(ctx.plan, lam_final)
                g = ctx.problem.contract(ctx.solver, ctx.plan, ab, ctx.theta)
                return torch.as_tensor(g, dtype=ctx.theta_dtype, device=ctx.theta_device), None

        cls._fn = _Fn
        return _Fn

    @classmethod
    def apply(cls, theta, problem):
        return cls.function().apply(theta, problem)
