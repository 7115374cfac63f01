"""Benchmark of the MI355X ETD2 propagation path.

Workloads (synthetic model; FP64 throughout):
  plain   : nuclear cascade, all transported isotopes A <= 56 (19520 states),
            CMB + two-blackbody EBL, pure Si-28 source, z 3 -> 0, dz 1e-3,
            one solve per launch sequence (K = 1).
  batched : the same operator with K = 512 source configurations (spectral
            index, rigidity cutoff, H/He/N/Si/Fe fractions, (1+z)^m evolution
            drawn from a seeded generator), one fold and one multi-RHS stage
            pair per step.

One benchmark "step" is one batched solve (K instances, 3000 ETD2 steps)
with every input already resident in HBM (the host-side setup happens
before the timed region).  The JSON line reports batched solves/s as
``value``, the plain solves/s beside it, the stage-kernel roofline measured
with HIP events on the launch stream, and the CPU oracle timed on a bounded
sample of the batched workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = (
    "solves/sec at A≤56 (batched vs plain, rel err≤1e-8); adjoint/forward ratio"
)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3, help="timed batched solves")
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--batch", type=int, default=512)
    p.add_argument("--plain-solves", type=int, default=3)
    p.add_argument("--cpu-sample-steps", type=int, default=12)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-events", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--adjoint-batch", type=int, default=64)
    p.add_argument("--no-uneven", action="store_true")
    p.add_argument("--no-em", action="store_true")
    p.add_argument("--no-uneven-em", action="store_true")
    p.add_argument("--uneven-em-batch", type=int, default=1024)
    p.add_argument("--em-batch", type=int, default=512)
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--uneven-batch", type=int, default=1024)
    p.add_argument("--uneven-steps", type=int, default=None,
                   help="march only the first N global steps of the uneven legs (profiler runs; "
                        "no column check)")
    p.add_argument("--z-max", type=float, default=3.0,
                   help="initial redshift (3.0 is the benchmark; smaller values shorten profiler runs)")
    p.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "etd2_stage_K512_traffic.json"),
                   help="rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the stage kernels (JSON)")
    return p.parse_args()


def measured_traffic(path, stage_groups):
    """HBM bytes per stage launch from a committed rocprofv3 counter summary:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each divided by what the same
    counter read per byte for an elementwise copy of known size, summed over
    the kernels of one stage and averaged over the two stages."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    per = []
    for names in stage_groups:
        tot = 0.0
        for n in names:
            ks = [v for k, v in rec["kernels"].items() if n in k]
            if len(ks) != 1:
                return None
            v = ks[0]
            tot += 1024.0 * (v["FETCH_SIZE_KiB"] / rec["fetch_factor"]
                             + v["WRITE_SIZE_KiB"] / rec["write_factor"])
        per.append(tot)
    return float(np.mean(per))


def timed(fn, torch, reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def stage_event_times(solver, plan, torch, lib_mod):
    """Average duration (s) of each stage kernel over one full solve, HIP events."""
    import ctypes

    lib = lib_mod.load()
    stream = torch.cuda.current_stream()
    nsteps = plan["nsteps"]
    plan["x"].zero_()
    solver.fold_windows(plan)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * nsteps + 1)]
    ev[0].record(stream)
    for k in range(nsteps):
        for s in (1, 2):
            lib_mod.check(
                lib.prh_etd2_stage(ctypes.byref(plan["op"]), ctypes.byref(plan["sch"]),
                                   ctypes.byref(plan["st"]), k, s, plan["stream"]),
                "stage",
            )
            ev[2 * k + s].record(stream)
    torch.cuda.synchronize()
    t1 = [ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(nsteps)]
    t2 = [ev[2 * k + 1].elapsed_time(ev[2 * k + 2]) for k in range(nsteps)]
    return float(np.mean(t1)) * 1e-3, float(np.mean(t2)) * 1e-3


def step_bytes_algorithmic(nnz_off, n, K):
    """Algorithmic bytes of one ETD2 step for K right-hand sides in a fused
    two-stage design: the off-diagonal CSR (8 B value + 4 B column) and the
    shared diagonal vectors once per stage, and per right-hand side and stage
    the state read, b read, a / F(x) or x written (4 x 8 B):
    2 (12 Z + 4 (n + 1)) + 64 n K + 64 n."""
    return 2.0 * (12.0 * nnz_off + 4.0 * (n + 1)) + 64.0 * n * K + 64.0 * n


def stage_bytes(plan, n, K):
    """Algorithmic HBM bytes of one stage launch of this implementation (x/a/fphi
    streams + block-band values + sources; the source term is formed from the
    injection grids, so no b vector is streamed)."""
    lay = plan["layout"]
    state = 24.0 * n * K
    matrix = 8.0 * (lay.fwd.zslots + n) + 16.0 * lay.blocks.shape[0]
    src = plan["bufs"]["src_grid"].numel() * 8.0
    return state + matrix + src


def _cfg_str(cfg):
    """Readable form of an autotuner configuration."""
    parts = []
    for k, v in cfg.items():
        parts.append(f"{k if isinstance(k, str) else k[0]}={v}")
    return ",".join(parts)


def stage_kernel_names(plan):
    """Per stage (1, 2): the names (as rocprofv3 prints them) of the kernels
    one stage launch of the plan runs."""
    kt, two = plan["kt"], "true" if plan["two"] else "false"
    wgw = plan["op"].wg_waves or 16
    dm = {None: 0, "pass": 1, "chunks": 2}[plan.get("dense_mode")]
    out = []
    for s2 in ("false", "true"):
        names = [f"etd2_stage<{s2}, {kt}, {two}, {wgw}, {dm}>"]
        if dm == 2:
            names.append(f"dense_finish<{s2}, {kt}, {wgw}>")
        out.append(names)
    return out


def fold_flops(plan):
    return 2.0 * plan["fold_rows"] * plan["dph"] * plan["nwin"]


def adjoint_leg(K, torch, log):
    """Adjoint/forward cost on the nuclear run, K instances, z 3 -> 0 (device time).

    forward  = rate fold + 3000 ETD2 steps keeping a checkpoint every 64 steps;
    backward = loss gradient w.r.t. the final state + recompute of every chunk +
               the adjoint sweep (two transposed stage kernels per step).
    Directional central differences (Richardson, relative steps 1e-5 and 1e-6)
    check the gradient.
    """
    from prince_amd.adjoint import AdjointProblem, draw_params

    prob = AdjointProblem(K=K, z_max=3.0)
    theta = draw_params(K, seed=7)
    prob.target()
    solver = prob._solver(theta)
    plan = solver.prepare(1e-3, with_adjoint=True)

    def fwd():
        solver.forward_checkpointed(plan)

    fwd()
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / reps
    st = plan["x"].permute(0, 2, 1)

    def bwd():
        stv = st.detach().clone().requires_grad_(True)
        losses, lp = prob.loss_of_state(stv)
        losses.sum().backward()
        ab = solver.adjoint_sweep(plan, stv.grad.permute(0, 2, 1).contiguous())
        return losses, lp, ab

    bwd()
    fwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses, lp, ab = bwd()
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    grad = prob.contract(solver, plan, ab, theta)
    # directional check: d/de L(theta + e v) vs grad . v, Richardson-extrapolated
    # central differences at two relative steps; per instance the step whose
    # h vs h/2 estimates agree best is kept (the loss takes log10 of fluxes
    # that span many decades, so a single step is either truncation- or
    # round-off-limited for some instances over 3000 steps)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(theta.shape)
    scale = np.maximum(np.abs(theta), 1.0)
    scale[:, 1] = 1.0 / np.log(10.0)

    def loss_at(t):
        return prob.loss_of_state(prob.forward(t, with_adjoint=False)[2], *lp)[0].detach().cpu().numpy()

    def cd(h):
        return (loss_at(theta + h * v * scale) - loss_at(theta - h * v * scale)) / (2 * h)

    an = (grad * v * scale).sum(axis=1)
    best_fd = np.full(K, np.nan)
    best_spread = np.full(K, np.inf)
    for h in (1e-5, 1e-6):
        c1, c2 = cd(h), cd(h / 2)
        rich = (4 * c2 - c1) / 3
        spread = np.abs(c2 - c1) / np.maximum(np.abs(c2), 1e-300)
        take = np.isfinite(rich) & (spread < best_spread)
        best_fd[take], best_spread[take] = rich[take], spread[take]
    # instances whose frozen interpolation knots change sign under the step give
    # NaN differences (the loss is not smooth there); they are left out
    ok = np.isfinite(best_fd)
    err = np.abs(an[ok] - best_fd[ok]) / np.maximum(np.abs(best_fd[ok]), 1e-300)
    rel = float(err.max()) if ok.any() else None
    conv = best_spread[ok] < 1e-3
    rel_conv = float(err[conv].max()) if conv.any() else None
    log(f"adjoint K={K}: forward {t_f*1e3:.1f} ms, backward {t_b*1e3:.1f} ms, "
        f"ratio {t_b/t_f:.2f}, directional FD rel dev max {rel} median "
        f"{float(np.median(err)) if ok.any() else None} on {int(ok.sum())}/{K} instances; "
        f"{int(conv.sum())} with converged differences: max {rel_conv}; "
        f"{int((err <= 1e-5).sum()) if ok.any() else 0} within 1e-5")
    return {
        "K": K,
        "forward_ms": round(t_f * 1e3, 2),
        "backward_ms": round(t_b * 1e3, 2),
        "ratio": round(t_b / t_f, 3),
        "fd_directional_max_rel_dev": rel,
        "fd_directional_median_rel_dev": float(np.median(err)) if ok.any() else None,
        "fd_directional_instances": int(ok.sum()),
        "fd_converged_instances": int(conv.sum()),
        "fd_converged_max_rel_dev": rel_conv,
        "fd_instances_within_1e-5": int((err <= 1e-5).sum()) if ok.any() else 0,
        "grad_finite": bool(np.all(np.isfinite(grad))),
        "params_per_instance": 8,
        "checkpoint_every": 64,
    }


def uneven_leg(run, K, torch, log, seed=0, steps=None):
    """Uneven batch (K columns, z_max ~ U[0.1, 5], EBL scale ~ U[0.6, 1.4], the
    batched source draws): one timed device solve after a 128-step warm-up, and
    two columns checked against a plain single-column solve of the same field."""
    from prince_amd import workloads
    from prince_amd.solvers.propagation import MultiRHSPropagationSolverETD2
    from prince_amd.solvers.uneven import UnevenBatchSolverETD2

    rng = np.random.default_rng(seed + 1)
    z_max = rng.uniform(0.1, 5.0, size=K)
    scale = rng.uniform(0.6, 1.4, size=K)
    srcs = workloads.batch_sources(run, K, seed=seed)
    g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    t0 = time.perf_counter()
    plan = g.prepare(1e-3)
    t_setup = time.perf_counter() - t0
    g.execute(plan, 0, min(128, plan["nsteps"]))  # warm-up; captures the chunk graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.execute(plan, 0, steps)
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t0
    nspec, d = run.spec_man.nspec, run.cr_grid.d
    host = plan["x"].cpu().numpy().transpose(0, 2, 1).reshape(nspec * d, K)
    state = np.empty_like(host)
    state[:, plan["order"]] = host
    devs = []
    for k in (int(np.argmax(z_max)), int(np.argmin(z_max))) if steps is None else ():
        one = MultiRHSPropagationSolverETD2(float(z_max[k]), 0.0, run, K=1)
        one.set_field_scales(1, [scale[k]])
        one[0].add_source_class(srcs[k])
        one.solve(1e-3)
        ref = np.asarray(one.state).reshape(-1)
        m = np.abs(ref) > 1e-30 * np.abs(ref).max()
        devs.append(float(np.max(np.abs(state[m, k] - ref[m]) / np.abs(ref[m]))))
    csteps = int(plan["col_steps"].sum())
    check = max(devs) if devs else None
    log(f"uneven K={K}: {plan['nsteps']} steps (column total {csteps}), {plan['nwin']} windows, "
        f"setup {t_setup:.1f}s, device {t_run*1e3:.0f} ms"
        + (f" -> {K/t_run:.1f} solves/s; column check {check:.2e}" if steps is None
           else f" for the first {steps} steps"))
    return {
        "K": K, "z_max_range": [0.1, 5.0], "ebl_scale_range": [0.6, 1.4],
        "global_steps": plan["nsteps"], "column_steps_total": csteps, "windows": plan["nwin"],
        "refold_group_steps": plan["refold_group"],
        "steps_timed": steps if steps is not None else plan["nsteps"],
        "setup_host_s": round(t_setup, 2), "ms_per_batch": round(t_run * 1e3, 1),
        "solves_per_s": round(K / t_run, 2) if steps is None else None,
        "column_steps_per_s": round(csteps / t_run, 1) if steps is None else None,
        "column_max_rel_dev_vs_two_field_solve": check,
    }


def uneven_em_leg(K, torch, log, seed=0, steps=None):
    """The uneven batch with the co-evolved EM cascade (K columns, z_max ~
    U[0.1, 5], EBL scale ~ U[0.6, 1.4]): the cascade transfer and Bethe-Heitler
    shapes of every column window are built on the device before the march.
    Reports the host setup, the device build of the window operators and one
    timed device march; two columns are checked against the same columns run
    as one-column batches."""
    from prince_amd import workloads
    from prince_amd.solvers.uneven import UnevenBatchSolverETD2

    run = workloads.nuclear_em_run()
    rng = np.random.default_rng(seed + 1)
    z_max = rng.uniform(0.1, 5.0, size=K)
    scale = rng.uniform(0.6, 1.4, size=K)
    srcs = workloads.batch_sources(run, K, seed=seed)
    g = UnevenBatchSolverETD2(z_max, run, ebl_scale=scale)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    t0 = time.perf_counter()
    plan = g.prepare(1e-3)
    t_setup = time.perf_counter() - t0
    t_build = g.timings.get("em_build_s", 0.0)
    g.execute(plan, 0, min(128, plan["nsteps"]))  # warm-up; captures the chunk graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.execute(plan, 0, steps)
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t0
    iters = plan["em_bufs"]["iters"].cpu().numpy()
    nspec, d, de = run.spec_man.nspec, run.cr_grid.d, run.em_grid.d
    host = plan["x"].cpu().numpy().transpose(0, 2, 1).reshape(nspec * d, K)
    gam = plan["em_bufs"]["gam"].cpu().numpy()
    state = np.empty((nspec * d + de, K))
    state[:, plan["order"]] = np.concatenate([host, gam], axis=0)
    devs = []
    for k in (int(np.argmin(z_max)), int(np.argsort(z_max)[K // 2])) if steps is None else ():
        one = UnevenBatchSolverETD2(z_max[k:k + 1], run, ebl_scale=scale[k:k + 1])
        one[0].add_source_class(srcs[k])
        one.solve(1e-3)
        ref = np.concatenate([one.state[:nspec * d, 0], one.state[-de:, 0]])
        for part in (slice(0, nspec * d), slice(nspec * d, None)):
            r, t = ref[part], state[part, k]
            m = np.abs(r) > 1e-30 * np.abs(r).max()
            devs.append(float(np.max(np.abs(t[m] - r[m]) / np.abs(r[m]))))
    csteps = int(plan["col_steps"].sum())
    del g, plan
    torch.cuda.empty_cache()
    check = max(devs) if devs else None
    log(f"uneven+EM K={K}: {csteps} column steps, {len(iters)} windows; setup {t_setup:.1f}s of which "
        f"device window operators {t_build:.1f}s; device march {t_run*1e3:.0f} ms"
        + (f" -> {K/t_run:.1f} solves/s ({K/(t_run + t_build):.1f} with the operator builds); "
           f"column check {check:.2e}" if steps is None else f" for the first {steps} steps"))
    if steps is not None:
        return {"K": K, "steps_timed": steps, "ms": round(t_run * 1e3, 1),
                "em_operator_build_s": round(t_build, 2)}
    return {
        "K": K, "z_max_range": [0.1, 5.0], "ebl_scale_range": [0.6, 1.4],
        "windows": int(len(iters)), "column_steps_total": csteps,
        "setup_s": round(t_setup, 2), "em_operator_build_s": round(t_build, 2),
        "em_operator_builds_per_s": round(len(iters) / t_build, 1) if t_build > 0 else None,
        "cascade_iterations_mean": float(np.mean(iters)), "cascade_iterations_max": int(iters.max()),
        "ms_per_batch": round(t_run * 1e3, 1), "solves_per_s": round(K / t_run, 2),
        "solves_per_s_with_operator_builds": round(K / (t_run + t_build), 2),
        "column_max_rel_dev_vs_one_column_batch": check,
    }


def em_leg(K, torch, log, seed=0):
    """Nuclear run with the co-evolved EM cascade (photons, e+-), K draws, z 3 -> 0:
    host setup (cascade transfers of the 288 windows, Bethe-Heitler shapes) and one
    timed device solve; the photon spectrum's energy is reported."""
    from prince_amd import workloads
    from prince_amd.solvers.propagation import MultiRHSPropagationSolverETD2

    run = workloads.nuclear_em_run()
    srcs = workloads.batch_sources(run, K, seed=seed)
    g = MultiRHSPropagationSolverETD2(3.0, 0.0, run, K=K)
    for k in range(K):
        g[k].add_source_class(srcs[k])
    t0 = time.perf_counter()
    plan = g.prepare(1e-3)
    t_setup = time.perf_counter() - t0
    g.execute(plan)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.execute(plan)
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t0
    gam = g.em_host_state(plan)[2 * run.em_grid.d:]
    E = run.em_grid.grid
    e_ph = float(np.mean((E * np.gradient(E))[:, None] * gam, axis=0).sum())
    # two columns against one-column solves of the plain path (nuclear and photons)
    from prince_amd.solvers.propagation import UHECRPropagationSolverETD2

    n, de = run.dim_states, run.em_grid.d
    full = g._finalize_state(plan["x"])
    devs = []
    for k in (0, K - 1):
        one = UHECRPropagationSolverETD2(3.0, 0.0, run)
        one.add_source_class(srcs[k])
        one.solve(1e-3)
        for part in (slice(0, n), slice(n + 2 * de, n + 3 * de)):
            r, t = one.state[part], full[part, k]
            m = np.abs(r) > 1e-30 * np.abs(r).max()
            devs.append(float(np.max(np.abs(t[m] - r[m]) / np.abs(r[m]))))
    log(f"EM cascade K={K}: setup {t_setup:.1f}s (cascade transfers for {plan['nwin']} windows "
        f"built on the device), device {t_run*1e3:.0f} ms -> {K/t_run:.1f} solves/s; "
        f"column check vs plain path {max(devs):.2e}")
    return {"K": K, "windows": plan["nwin"], "setup_s": round(t_setup, 2),
            "ms_per_batch": round(t_run * 1e3, 1), "solves_per_s": round(K / t_run, 2),
            "mean_photon_energy_density": e_ph, "cascade_iterations_max": int(max(g.em_iterations)),
            "column_max_rel_dev_vs_plain": max(devs)}


def host_cpu():
    """(threads available to this process, CPU model name)."""
    try:
        nthreads = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        nthreads = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return nthreads, model


def cpu_baseline(run, srcs, zmax, nsteps, K, args, log):
    """The oracle (scipy CSR SpMV single-threaded as the reference's scipy
    backend, the dense fold through numpy/BLAS on every host thread): full
    plain A <= 56 solves (median of up to 3 after a warm-up) -- the batch of K
    is K such sequential solves -- and a K-column multi-RHS sample whose steps
    after the first window are extrapolated to the full schedule."""
    from threadpoolctl import threadpool_limits

    from oracle.etd2_cpu import OracleETD2
    from prince_amd import workloads

    nthreads, model = host_cpu()
    with threadpool_limits(limits=nthreads):
        times = []
        for rep in range(4):
            o = OracleETD2(zmax, 0.0, run)
            o.add_source_class(workloads.silicon_source(run))
            t0 = time.perf_counter()
            o.solve(1e-3)
            times.append(time.perf_counter() - t0)
            if rep == 0 and times[0] > 40.0:  # slow host: one timed solve after the warm-up
                o.solve(1e-3)
                times.append(time.perf_counter() - t0 - times[0])
                break
        t_plain = float(np.median(times[1:]))
        # multi-RHS sample: first window (split and index maps) timed apart
        ns = args.cpu_sample_steps
        om = OracleETD2(zmax, zmax - ns * 1e-3, run, K=K)
        for k in range(K):
            om.add_source_class(srcs[k], k)
        marks = []
        t0 = time.perf_counter()
        om.solve(1e-3, step_callback=lambda k_, s_: marks.append(time.perf_counter()))
        first = marks[0] - t0
        per_step = (marks[-1] - marks[0]) / (len(marks) - 1)
    t_multi = first + per_step * (nsteps - 1)
    log(f"cpu oracle ({nthreads} threads, {model}): plain solve {t_plain:.2f} s (median of "
        f"{len(times) - 1}); K={K} multi-RHS sample {len(marks)} steps -> {t_multi:.1f} s per batch")
    return {
        "value": round(1.0 / t_plain, 4),
        "unit": "solves/s",
        "cores": nthreads,
        "kind": "port",
        "cpu_model": model,
        "sample": f"FP64 oracle, full plain A<=56 solves z {zmax:g}->0 ({nsteps} steps), median of "
        f"{len(times) - 1} after a warm-up, dense fold on {nthreads} BLAS threads, CSR SpMV "
        f"single-threaded; the K={K} batch counts as {K} sequential solves",
        "plain_solve_s": round(t_plain, 3),
        "multi_rhs_solves_per_s": round(K / t_multi, 4),
        "multi_rhs_sample": f"K={K} columns, {len(marks)} of {nsteps} steps: first step "
        f"(first-window fold, split and index maps) {first:.2f} s, then {per_step*1e3:.0f} ms per "
        "step, extrapolated by step count",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the ETD2 path shards by instance with no per-step "
            "exchange, and this benchmark runs and reports the single-GPU share only.",
            file=sys.stderr,
        )
        sys.exit(2)
    import torch

    from prince_amd import _lib, workloads
    from prince_amd.solvers.propagation import (
        MultiRHSPropagationSolverETD2,
        UHECRPropagationSolverETD2,
    )

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    # calibration copy of a known size (256 MiB read + 256 MiB written) for
    # scaling the profiler's FETCH_SIZE / WRITE_SIZE counters
    calib_src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    calib_dst = torch.empty_like(calib_src)
    torch.mul(calib_src, 1.0, out=calib_dst)  # one elementwise kernel: 256 MiB in, 256 MiB out
    torch.cuda.synchronize()
    del calib_src, calib_dst
    t0 = time.perf_counter()
    run = workloads.nuclear_run()
    n = run.dim_states
    log(f"model built in {time.perf_counter() - t0:.1f}s (n={n})")

    # ---------------- plain path (K = 1)
    zmax = args.z_max
    plain = UHECRPropagationSolverETD2(zmax, 0.0, run)
    plain.add_source_class(workloads.silicon_source(run))
    pplan = plain.prepare(1e-3)
    plain.execute(pplan)
    torch.cuda.synchronize()
    t_plain = timed(lambda: plain.execute(pplan), torch, args.plain_solves) / args.plain_solves
    plain_state = plain._finalize_state(pplan["x"])
    nsteps = pplan["nsteps"]
    log(f"plain: {t_plain * 1e3:.2f} ms/solve ({nsteps} steps, {pplan['nwin']} windows)")

    # ---------------- batched path (K = batch)
    K = args.batch
    multi = MultiRHSPropagationSolverETD2(zmax, 0.0, run, K=K)
    srcs = workloads.batch_sources(run, K, seed=0)
    for k in range(K):
        multi[k].add_source_class(srcs[k])
    t0 = time.perf_counter()
    tuning = None
    if not args.no_autotune:
        table = multi.autotune(1e-3, steps=30)
        tuning = {
            "seconds": round(time.perf_counter() - t0, 1),
            "chosen": _cfg_str(table[0][1]),
            "table_us_per_step": [[round(t * 1e6, 1), _cfg_str(c)] for t, c in table],
        }
        log(f"autotune ({tuning['seconds']}s): {tuning['chosen']} "
            f"{table[0][0]*1e6:.1f} us/step; slowest {table[-1][0]*1e6:.1f}")
    t0 = time.perf_counter()
    bplan = multi.prepare(1e-3)
    log(f"batched plan K={K} prepared in {time.perf_counter() - t0:.1f}s")
    for _ in range(max(0, args.warmup)):
        multi.execute(bplan)
    torch.cuda.synchronize()
    t_batch = timed(lambda: multi.execute(bplan), torch, args.steps)
    ms_per = t_batch / args.steps * 1e3
    batched_sps = K * args.steps / t_batch
    plain_sps = 1.0 / t_plain
    log(f"batched: {ms_per:.1f} ms/solve-batch -> {batched_sps:.1f} solves/s")

    # ---------------- end to end: solve() including the host setup (prepare)
    t0 = time.perf_counter()
    multi.solve(1e-3)
    t_e2e = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain.solve(1e-3)
    t_e2e_plain = time.perf_counter() - t0
    log(f"end to end solve(): batched {t_e2e:.2f} s -> {K / t_e2e:.1f} solves/s; "
        f"plain {t_e2e_plain:.3f} s")
    bplan = multi._plan

    # ---------------- parity of the batched columns vs the plain path
    bstate = multi._finalize_state(bplan["x"])
    check_cols = [0, K // 3, K - 1]
    col_dev = []
    for k in check_cols:
        one = UHECRPropagationSolverETD2(zmax, 0.0, run)
        one.add_source_class(srcs[k])
        one.solve(1e-3)
        ref = one.state
        peak = np.abs(ref).max()
        mask = np.abs(ref) > 1e-30 * peak
        col_dev.append(float(np.max(np.abs(bstate[mask, k] - ref[mask]) / np.abs(ref[mask]))))
    log(f"batched-vs-plain column deviation: {max(col_dev):.2e}")

    # ---------------- roofline of the dominant kernel (stage kernels, batched)
    roof = None
    kernels = {}
    if not args.no_events:
        t1s, t2s = stage_event_times(multi, bplan, torch, _lib)
        # roofline: the algorithmic bytes of a fused two-stage step (half a step
        # per stage launch), and beside it this implementation's own count
        nbytes = 0.5 * step_bytes_algorithmic(bplan["layout"].nnz_off, n, K)
        own = stage_bytes(bplan, n, K)
        avg = 0.5 * (t1s + t2s)
        achieved = nbytes / avg / 1e9
        roof = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.traffic, stage_kernel_names(bplan)),
            "kernel": "etd2_stage (stage 1 and 2, batched K=%d)" % K,
            "bytes_per_launch": nbytes,
            "avg_launch_us": round(avg * 1e6, 2),
            "implementation_bytes_per_launch": own,
            "implementation_frac": round(own / avg / 1e9 / HBM_PEAK_GBS, 4),
        }
        kernels["etd2_stage1_us"] = round(t1s * 1e6, 2)
        kernels["etd2_stage2_us"] = round(t2s * 1e6, 2)
        # fold GEMM (MFMA) timing
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        reps = 5
        for _ in range(reps):
            multi.fold_windows(bplan)
        ev1.record()
        torch.cuda.synchronize()
        tf = ev0.elapsed_time(ev1) * 1e-3 / reps
        kernels["fold_gemm_us"] = round(tf * 1e6, 1)
        kernels["fold_gemm_tflops"] = round(fold_flops(bplan) / tf / 1e12, 2)
        kernels["fold_gemm_frac_fp64_mfma"] = round(fold_flops(bplan) / tf / 78.6e12, 4)
        log(f"stage1 {t1s*1e6:.1f} us, stage2 {t2s*1e6:.1f} us, {achieved:.0f} GB/s; "
            f"fold {tf*1e6:.0f} us")

    # ---------------- adjoint (C4 geometry): backward / forward device time
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args.adjoint_batch, torch, log)

    # ---------------- uneven batch (per-column z_max and EBL scale)
    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(run, args.uneven_batch, torch, log, steps=args.uneven_steps)

    # ---------------- EM cascade (shared schedule)
    em = None
    if not args.no_em:
        em = em_leg(args.em_batch, torch, log)

    # ---------------- uneven batch with the EM cascade (C3 as specified)
    uneven_em = None
    if not args.no_uneven_em:
        uneven_em = uneven_em_leg(args.uneven_em_batch, torch, log, steps=args.uneven_steps)

    # ---------------- CPU baseline: the FP64 oracle on the host cores
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(run, srcs, zmax, nsteps, K, args, log)

    out = {
        "metric": METRIC,
        "value": round(batched_sps, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"nuclear_batch: A<=56 ({run.spec_man.nspec} species, {n} states), "
            f"CMB+EBL, K={K} seeded source draws, z {zmax:g}->0, dz=1e-3",
            "K": K,
            "nsteps": nsteps,
            "windows": pplan["nwin"],
            "nnz_offdiag": int(pplan["layout"].nnz_off),
        },
        "plain_solves_per_s": round(plain_sps, 3),
        "plain_ms_per_solve": round(t_plain * 1e3, 3),
        "batched_vs_plain": round(batched_sps / plain_sps, 2),
        "steps_per_s_batched": round(nsteps * args.steps / t_batch, 1),
        "steps_per_s_plain": round(nsteps / t_plain, 1),
        "batched_column_max_rel_dev_vs_plain": max(col_dev),
        "end_to_end_solve_s_batched": round(t_e2e, 3),
        "end_to_end_solves_per_s_batched": round(K / t_e2e, 2),
        "end_to_end_solve_s_plain": round(t_e2e_plain, 4),
        "adjoint_forward_ratio": adjoint["ratio"] if adjoint else None,
        "adjoint": adjoint,
        "uneven": uneven,
        "em_cascade": em,
        "uneven_em": uneven_em,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "autotune": tuning,
    }
    _ = plain_state
    print(json.dumps(out))


if __name__ == "__main__":
    main()
