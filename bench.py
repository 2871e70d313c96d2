#!/usr/bin/env python
"""Benchmark of the lattice-Boltzmann elasticity / Darcy / Biot path on one MI355X.

Workload of the headline line: batched poroelasticity, 64 independent periodic
Biot instances with square grids drawn from {64, 128, 256, 512} cells per side,
Lame parameters lambda, mu ~ U[0.5, 2] and log-permeability mean ~ U[-1, 1] with
seeded log-normal heterogeneity (variance 1, correlation length 16 cells), FP64;
in every physical step each instance's elasticity sub-problem is iterated with
multigrid V-cycles (8 smoothing steps per level, per-instance depth chosen by
the depth tuner) to a residual of 1e-8 of the residual of the zero state.  One
"step" is one physical time step of all 64 instances.  The instances run as one
batch (batched launches over every instance's tiles, captured V-cycles, device-
side per-instance convergence masks) and, for the comparison, one after another
with their own captured V-cycles (the plain solves); both runs do the same
physical steps and are checked to agree bit for bit.  ``value`` is MLUPS of the
batched run in the reference's convention for the Biot solver (fine cells x
physical steps / s / 1e6), with all inputs resident in HBM.

``roofline`` prices the dominant kernel of the batched cycle, the level-0
smoothing launch of all instances on LDS tiles (two damped smoothing steps per
launch with the fused coupling force), at its algorithmic bytes per launch (per
cell 8 populations read, the force's 2 components read, 8 populations and the
divergence written: 19 values x 8 B) over its average launch duration with every
instance active, timed with HIP events on the stream it is launched on.
``cpu_baseline`` times the CPU restatement of the reference algorithm (torch
CPU, float64, all host cores) on a bounded sample of the same workload: the first
timed physical step of instances of the set, each with the number of V-cycles
its GPU solve needed.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "xlb-biot_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "MLUPS & HBM %peak, plain vs batched; adjoint/forward time at residual 1e-8"
HBM_PEAK_GBS = 8000.0
# the dominant kernel of the batched-Biot headline (level-0 smoothing of every
# instance on LDS tiles: periodic grids, no residual input, fused coupling force)
C5_KERNEL = "k_tile_smooth_batch<double, 0, false, true, false>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def event_pair_time(torch, fn, reps):
    """Average duration (s) of ``fn()`` launches, each bracketed by its own HIP events."""
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for i in range(reps):
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    return sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / reps / 1e3


def darcy_workload(n, precision):
    import problems as P
    return P.darcy_c2("gpu", n=n, precision=precision, heterogeneous=True, seed=0, variance=1.0, corr=16.0)


def bench_darcy(torch, n, precision, reps, warmup=2, chunk=64, tune=True, tile_steps=0, tile_threads=0):
    """Plain Darcy steps of the C2 set-up on an n x n domain (+ ghost layers), replayed
    from the step loop's captured graph (``chunk`` steps, only launches of the tuned
    kind: one- or two-step sweeps or m-step tiles).  ``reps`` whole replays are timed,
    wall clock and HIP events on the stream they run on; the state stays in the graph's
    own buffers, so no copy is in the timed region.  The average launch duration is
    the replays' GPU time over the launches they hold; a launch reads and writes every
    population once (19 values per cell with the per-cell rate)."""
    from src import autotune
    d = darcy_workload(n, precision)
    if tune:
        d.autotune()
    if tile_steps:
        d.step_launch = autotune.tile_choice(tile_steps, tile_threads)
    choice, spl, _ = d._multi()
    d.run_steps(d.assign_fields_sharded(), 1, chunk)  # builds the runner: buffers and the captured graph
    runner = d._runners[("steps", chunk, choice)]
    if runner.graph is None:
        raise RuntimeError(f"chunk {chunk} holds no whole graph of {spl}-step launches")
    for _ in range(warmup):
        runner.graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(reps):
        runner.graph.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu = e0.elapsed_time(e1) / 1e3
    steps = reps * runner.gsteps
    launches = reps * (runner.gsteps // spl)
    S = 8 if precision == "f64/f64" else 4
    cells = d.nx * d.ny
    kind = "tile" if choice > autotune.TILE_BASE else ("sweep2" if choice == autotune.SWEEP2 else "sweep")
    threads, blocks = None, None
    if kind == "tile":
        threads = autotune.tile_threads(choice) or autotune.library_tile_threads(d.nx, d.ny, spl)
        blocks = autotune.tile_blocks(d.nx, d.ny, spl, threads)
    return dict(elapsed=wall, gpu=gpu, steps=steps, cells=cells, kernel_s=gpu / launches, bytes=cells * 19 * S,
                steps_per_launch=spl, launch_kind=kind, tile_threads=threads, tile_blocks=blocks, solver=d,
                graph_steps=runner.gsteps)


def c2_record(torch, n=512, reps=40):
    """C2 plain Darcy, FP64 and FP32: MLUPS over whole graph replays, the average
    duration and compulsory-byte rate of the launched kernel, measured traffic where
    the kernel and launch grid were counter-profiled."""
    from src import autotune
    out = {}
    for precision in ("f64/f64", "f32/f32"):
        r = bench_darcy(torch, n, precision, reps)
        d = r["solver"]
        kind = r["launch_kind"]
        T = "double" if precision == "f64/f64" else "float"
        if kind == "tile":
            tx_, ty_ = autotune.TILE_SHAPES[r["tile_threads"]]
            kname = f"k_tile_steps<{T}, xb::DarcyColl<{T}>, 2, {tx_}, {ty_}, {r['tile_threads']}>"
        else:
            kname = (f"k_sweep2<{T}, xb::DarcyColl<{T}>, 2>" if kind == "sweep2"
                     else f"k_sweep<{T}, xb::DarcyColl<{T}>, 0, 2, false, false>")
        gbs = r["bytes"] / r["kernel_s"] / 1e9
        rec = {"grid": [d.nx, d.ny], "mlups": round(r["cells"] * r["steps"] / r["elapsed"] / 1e6, 1),
               "timed_steps": r["steps"], "graph_steps": r["graph_steps"], "kernel": kname,
               "steps_per_launch": r["steps_per_launch"], "threads_per_block": r["tile_threads"],
               "launch_blocks": r["tile_blocks"], "kernel_us": round(r["kernel_s"] * 1e6, 2),
               "compulsory_bytes_per_launch": r["bytes"], "gbs": round(gbs, 1),
               "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
               "per_step_equivalent_gbs": round(gbs * r["steps_per_launch"], 1)}
        if kind == "tile":
            rec["traffic"] = profiled_traffic(kname, r["tile_blocks"], r["cells"])
        out["fp64" if precision == "f64/f64" else "fp32"] = rec
    return out


def bench_large_elasticity(torch, n, reps, warmup=1, precision="f64/f64", chunk=32):
    """Periodic D2Q8 steps on an n x n grid with a body force (18 values per cell),
    timed over whole replays of the step loop's captured graph (as bench_darcy)."""
    import problems as P
    from src import autotune
    el = P.periodic_elasticity("gpu", nx=n, precision=precision)
    el.autotune()
    choice, spl, _ = el._multi(None)
    el.run_steps(el.assign_fields_sharded(), 1, chunk)
    runner = el._runners[("steps", chunk, True, choice)]
    for _ in range(warmup):
        runner.graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(reps):
        runner.graph.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu = e0.elapsed_time(e1) / 1e3
    S = 8 if precision == "f64/f64" else 4
    threads = None
    if choice > autotune.TILE_BASE:
        threads = autotune.tile_threads(choice) or autotune.library_tile_threads(n, n, spl)
    return dict(elapsed=wall, gpu=gpu, steps=reps * runner.gsteps, cells=n * n,
                kernel_s=gpu / (reps * (runner.gsteps // spl)), bytes=n * n * 18 * S, steps_per_launch=spl,
                tile_threads=threads)


def c2_time_to_solution(torch, n=512, tol=1e-8, check_every=1000, checks=10, max_steps=4000000):
    """C2 (FP64) from rest to ``|f_{k+1} - f_k| < tol |f_1 - f_0|``, the test every
    ``check_every`` steps on the device (``run_steps_to``: captured blocks, one host
    read per ``checks`` tests); the launch shape is the tuned one."""
    d = darcy_workload(n, "f64/f64")
    d.autotune()
    f0 = d.assign_fields_sharded()
    # buffers, tuned launches and the captured blocks are built before the clock starts
    d.run_steps_to(f0, tol, check_every, max_steps=check_every * checks + 1, checks=checks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, k, ratio = d.run_steps_to(f0, tol, check_every, max_steps=max_steps, checks=checks)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    cells = d.nx * d.ny
    rec = {"grid": [d.nx, d.ny], "tol": tol, "steps": k, "seconds": round(sec, 4), "rel_residual": ratio,
           "check_every": check_every, "host_reads_every_steps": check_every * checks}
    if k > 0:
        rec["mlups"] = round(cells * k / sec / 1e6, 1)
    return rec


def bench_c1(torch, tol=1e-10, check_every=10, checks=10, max_steps=20000):
    """C1: periodic 64^2 elasticity, FP64, plain steps until the relative displacement
    residual ||f_{k+1}-f_k|| / ||f_1-f_0|| < tol (tested every ``check_every`` steps on
    the device, ``run_steps_to``); error to the closed form."""
    import numpy as np
    import problems as P
    el = P.periodic_elasticity("gpu", nx=64)
    f0 = el.assign_fields_sharded()
    # the runner (buffers, captured blocks) is built before the clock starts
    el.run_steps_to(f0, tol, check_every, max_steps=check_every * checks + 1, checks=checks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, k, ratio = el.run_steps_to(f0, tol, check_every, max_steps=max_steps, checks=checks)
    torch.cuda.synchronize()
    el_s = time.perf_counter() - t0
    eta, _ = el.update_macroscopic(f)
    ue = P.manufactured_c1(64) / (8 * np.pi ** 2)
    err = float(np.linalg.norm(eta.cpu().numpy() - ue) / np.linalg.norm(ue))
    return {"steps": k, "seconds": round(el_s, 5), "ms": round(el_s * 1e3, 3),
            "mlups": round(64 * 64 * k / el_s / 1e6, 1), "rel_residual": ratio, "rel_error_closed_form": round(err, 6),
            "check_every": check_every, "host_reads_every_steps": check_every * checks}


def bench_c3(torch, n=256, levels=4, steps=20, tol=1e-8, lam=1.0, mu=1.0, K=1.0, max_cycles=300):
    """Periodic Biot (manufactured forcing), FP64, elasticity V-cycles (k=2) to an inner
    relative residual ``tol`` in every physical step (best state kept when a step's
    residual stops improving for 20 cycles)."""
    import problems as P
    b = P.biot_periodic("gpu", nx=n, levels=levels, steps=1, lam=lam, mu=mu, K=K)
    b.elasticityTol = tol
    b.elasticityMaxSteps = max_cycles
    b.elasticityPatience = 20
    fe, fd, dde = b.initial_state()
    out = b._step_soa(fe, fd, dde, 0)  # warm-up step: captures the V-cycle graphs
    fe, fd, dde = out[0], out[1], out[4]
    b.inner_counts.clear()
    b.inner_residuals.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    diverged = None
    done = 0
    for t in range(1, steps + 1):
        try:
            out = b._step_soa(fe, fd, dde, t)
        except FloatingPointError as e:
            # the reference V-cycle diverges for some materials and depths: the inner
            # loop stops on a non-finite residual
            diverged = {"physical_step": t, "message": str(e)}
            break
        fe, fd, dde = out[0], out[1], out[4]
        done = t
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cycles = sum(b.inner_counts)
    rels = list(b.inner_residuals) or [float("nan")]
    if diverged is not None:
        return {"physical_steps": steps, "physical_steps_completed": done, "diverged": diverged,
                "seconds": round(el, 4), "vcycles": cycles,
                "steps_reaching_tol": sum(1 for r in rels if r <= tol)}
    # time of one captured V-cycle (with the nonlinear Biot force), HIP events on the stream
    vg = b._inner
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        vg.replay()
    e1.record()
    torch.cuda.synchronize()
    vc_s = e0.elapsed_time(e1) / 1e3 / 20
    # the same V-cycle with the coupling force formed by separate divergence and gradient kernels
    b.fusedCoupling = False
    b._step_soa(fe, fd, dde, steps + 1)
    vg2 = b._inner
    e0.record()
    for _ in range(20):
        vg2.replay()
    e1.record()
    torch.cuda.synchronize()
    vc2_s = e0.elapsed_time(e1) / 1e3 / 20
    reached = sum(1 for r in rels if r <= tol)
    return {"physical_steps": steps, "physical_steps_completed": done, "seconds": round(el, 4), "vcycles": cycles,
            "steps_reaching_tol": reached, "worst_rel_residual": max(rels),
            "median_rel_residual": sorted(rels)[len(rels) // 2],
            "vcycles_per_step": round(cycles / steps, 2), "mlups_reference_convention": round(n * n * steps / el / 1e6, 3),
            "vcycle_us": round(vc_s * 1e6, 2), "vcycle_us_unfused_coupling": round(vc2_s * 1e6, 2),
            "fine_cell_vcycles_per_s_millions": round(n * n / vc_s / 1e6, 1)}


def bench_copy(torch, reps=20):
    """Achievable HBM bandwidth: a plain 1 GiB device-to-device copy (read + write)."""
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    t = event_pair_time(torch, lambda: b.copy_(a), reps)
    gbs = 2 * a.numel() * 8 / t / 1e9
    return {"bytes_per_launch": 2 * a.numel() * 8, "us": round(t * 1e6, 1), "gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_biot_vcycle(torch, n=4096, levels=6, reps=10, eager=False):
    """One fused-coupling V-cycle of the periodic Biot step (manufactured forcing,
    lambda = mu = 1, K = 1, k = 2 smoothing steps, FP64) on an n x n grid, replayed
    from its captured graph (``eager``: launched one kernel at a time); bytes from
    :func:`vcycle_bytes`."""
    import problems as P
    b = P.biot_periodic("gpu", nx=n, levels=levels, steps=1, device_hooks=True)
    el = b.elasticity
    el.autotune(external_force=True)
    fe, fd, dde = b.initial_state()
    ctx = b._pre_inner(fe, fd, dde, 1)
    if eager:
        div = el._divergence_soa(fe)

        def cycle():
            el._vcycle_nl(fe, div, 0, None, ctx["force"], ctx["nlf"])
        cycle()
        t = event_pair_time(torch, cycle, reps)
    else:
        vg = el.vcycle_graph(ctx["force"], ctx["nlf"])
        vg.load(fe)
        vg.run(2)
        t = event_pair_time(torch, vg.replay, reps)
    m = el._smooth_tile_steps(0, ctx["nlf"])
    byt = vcycle_bytes(n, levels, el.k_smooth, m=m)
    byt1 = vcycle_bytes(n, levels, el.k_smooth)
    gbs, gbs1 = byt / t / 1e9, byt1 / t / 1e9
    return {"grid": n, "levels": levels, "k_smooth": el.k_smooth, "smoothing_steps_per_launch": m,
            "vcycle_ms": round(t * 1e3, 3),
            "compulsory_gb": round(byt / 1e9, 3), "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "single_sweep_formulation_gb": round(byt1 / 1e9, 3),
            "single_sweep_equivalent_gbs": round(gbs1, 1),
            "single_sweep_equivalent_hbm_frac": round(gbs1 / HBM_PEAK_GBS, 4),
            "fine_cell_vcycles_per_s_millions": round(n * n / t / 1e6, 1)}


def vcycle_bytes(n, levels, k, S=8, m=1):
    """Algorithmic HBM bytes of one fused-coupling V-cycle of a periodic n x n grid
    whose smoothing runs ``m`` steps per launch (``m = 1``: one sweep per step).  A
    smoothing sweep of level 0 reads f, the force and the divergence and writes f and
    the divergence (20 S per cell); a tile launch of m steps forms the divergence
    itself (19 S per launch); coarse levels also read the restricted residual (26 S,
    25 S).  The residual sweep reads f, force and divergence and writes the
    restriction (11 S + 8 S / 4; coarse levels 17 S + 8 S / 4).  With tile launches
    (m > 1, k > 1) a coarse level's first launch starts from the zero state without
    reading one (8 S less), and the prolongation rides in the first post-smoothing
    launch, which then also reads the coarse correction (8 S / 4 per fine cell);
    otherwise coarse states are zero-filled (9 S) and prolonged by a kernel of their
    own (8 S / 4 read coarse, 16 S + 1 S per fine cell)."""
    total = 0.0
    launches = 2 * (-(-k // m))
    for lvl in range(levels):
        c = (n >> lvl) ** 2
        coarsest = lvl == levels - 1
        if m == 1:
            smooth = 20 if lvl == 0 else 26
        else:
            smooth = 19 if lvl == 0 else 25
        fused = m > 1 and k > 1
        total += launches * smooth * c
        if lvl > 0 and fused:
            total -= 8 * c                                 # the zero start state is not read
        if not coarsest:
            total += ((11 if lvl == 0 else 17) + 2) * c   # residual + restriction out
            if not fused:
                total += 9 * c / 4                         # zero fill of the coarse state
            total += (2 if fused else 2 + 17) * c          # prolongation
    return total * S


def c5_build(torch, n_inst=64, tol=1e-8, seed=3, max_cycles=3000, k_smooth=8, sequential=True):
    """The batched-Biot set: ``n_inst`` periodic Biot instances (sizes from {64, 128,
    256, 512}, lambda, mu ~ U[0.5, 2], log-permeability mean ~ U[-1, 1], log-normal K),
    FP64, elasticity V-cycles to ``tol`` of the zero state's residual in every
    physical step.  Each instance's hierarchy depth is tuned first (shortest predicted
    time to the tolerance; cached) and used by both runs: the batch (batched launches,
    device-side convergence masks, captured batched cycles) and, with ``sequential``,
    the same instances one after another (each with its captured V-cycle graph and a
    host-side residual test per cycle).  The V-cycles smooth ``k_smooth`` = 8 times
    per level (the reference's MG_n_smoothing): with the default 2 (or 4) the
    reference V-cycle diverges for these Lame parameters."""
    import collections
    import problems as P
    from src import autotune
    from src.batch import BiotBatch
    specs = P.c5_specs(n_inst, seed)
    fields = [P.c5_permeability(s) for s in specs]
    t0 = time.perf_counter()
    levels = []
    for s, K in zip(specs, fields):
        n = s["n"]
        cands = [L for L in range(2, 6) if (n >> (L - 1)) >= 16]
        key = (f"mgdepth/biot_periodic/n{n}/lam{s['lam']:.6f}/mu{s['mu']:.6f}/logk{s['logk_mean']:.6f}"
               f"/seed{s['seed']}/k{k_smooth}/tol{tol:g}")
        L, _ = autotune.tune_levels(
            key, lambda L, s=s, K=K: P.biot_c5("gpu", s, L, tol, max_cycles, K=K, k_smooth=k_smooth).probe_levels(),
            cands, tol)
        levels.append(L)
    # launch shapes of the single-grid sweeps for every (size, depth) in the set
    for n, L in sorted(set((s["n"], L) for s, L in zip(specs, levels))):
        spec = next(s for s, L_ in zip(specs, levels) if s["n"] == n and L_ == L)
        P.biot_c5("gpu", spec, L, tol, max_cycles, k_smooth=k_smooth,
                  K=fields[specs.index(spec)]).elasticity.autotune(external_force=True)
    tune_s = time.perf_counter() - t0
    log(f"  depths and launch shapes tuned in {tune_s:.1f} s: {dict(collections.Counter(levels))}")
    seq = None
    if sequential:
        seq = [P.biot_c5("gpu", s, L, tol, max_cycles, K=K, k_smooth=k_smooth)
               for s, L, K in zip(specs, levels, fields)]
        for p in seq:
            p.prepare_inner()
    t0 = time.perf_counter()
    bat = BiotBatch([P.biot_c5("gpu", s, L, tol, max_cycles, K=K, k_smooth=k_smooth)
                     for s, L, K in zip(specs, levels, fields)])
    log(f"  batch built in {time.perf_counter() - t0:.1f} s, smoothing launch shapes {bat.batch.smooth_rows}")
    return dict(specs=specs, fields=fields, levels=levels, seq=seq, bat=bat, tune_s=tune_s, tol=tol,
                k_smooth=k_smooth, max_cycles=max_cycles)


def c5_dominant_plan(bat):
    """The batch's first level-0 smoothing launch: every instance's finest grid on LDS
    tiles, two damped smoothing steps with the fused coupling force."""
    for p in bat.batch.schedules[0]:
        if p.tag is not None and p.tag[:2] == ("smooth", 0):
            return p
    raise RuntimeError("no level-0 smoothing launch in the batched schedule")


def c5_kernel_time(torch, bat, reps):
    """Average duration of the dominant launch with every instance active: ``reps``
    launches between two HIP events on the stream they are launched on."""
    plan = c5_dominant_plan(bat)
    active = torch.ones(bat.batch.n, dtype=torch.int32, device="cuda")
    # with every instance active a solve's cycles launch one workgroup per tile (the
    # work-list form takes over once most of the work has converged)
    plan.use_list(False)
    plan.launch(active)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.launch(active)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps, plan


def bench_c5(torch, steps=10, warmup=2, n_inst=64, tol=1e-8, seed=3, max_cycles=3000, k_smooth=8, kernel_reps=20,
             build=None):
    """The headline: ``warmup`` untimed physical steps, then ``steps`` timed physical
    steps of every instance, batched and one after another (the same steps, from the
    same states; the final states are compared bit for bit); then the dominant
    kernel's launch duration with every instance active."""
    import collections
    b = build or c5_build(torch, n_inst, tol, seed, max_cycles, k_smooth)
    specs, levels, seq, bat = b["specs"], b["levels"], b["seq"], b["bat"]
    cells = sum(s["n"] ** 2 for s in specs)

    def seq_step(states, t):
        out = []
        for p, st in zip(seq, states):
            o = p._step_soa(*st, t)
            out.append((o[0], o[1], o[4]))
        return out

    def bat_step(states, t):
        return [(o[0], o[1], o[4]) for o in bat.step(states, t)]

    seq_states = [p.initial_state() for p in seq]
    bat_states = bat.initial_state()
    for t in range(warmup):
        bat_states = bat_step(bat_states, t)
        seq_states = seq_step(seq_states, t)
    # states entering the first timed step (for the CPU sample: the first instance of
    # every (size, depth) pair of the set)
    firsts = {}
    for i, (s, L) in enumerate(zip(specs, levels)):
        firsts.setdefault((s["n"], L), i)
    cpu_start = {i: tuple(x.detach().cpu() for x in seq_states[i]) for i in firsts.values()}

    def timed(fn, states):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for t in range(warmup, warmup + steps):
            states = fn(states, t)
        e1.record()
        torch.cuda.synchronize()
        return states, time.perf_counter() - t0, e0.elapsed_time(e1) / 1e3

    for k in bat.timing:
        bat.timing[k] = 0.0
    log(f"  batched: {steps} timed physical steps after {warmup}")
    bat_states, bat_s, bat_gpu = timed(bat_step, bat_states)
    log(f"  batched {bat_s:.3f} s; sequential: the same steps")
    seq_states, seq_s, seq_gpu = timed(seq_step, seq_states)
    log(f"  sequential {seq_s:.3f} s")
    same = all(all(torch.equal(x, y) for x, y in zip(a, c)) for a, c in zip(seq_states, bat_states))
    timed_counts = [p.inner_counts[warmup:warmup + steps] for p in seq]
    rels = [r for p in seq for r in p.inner_residuals[warmup:warmup + steps]]
    counts = [c for cs in timed_counts for c in cs]
    fine_cycles = sum(s["n"] ** 2 * sum(cs) for s, cs in zip(specs, timed_counts))
    kern_s, plan = c5_kernel_time(torch, bat, kernel_reps)
    return {"cells": cells, "batched_s": bat_s, "batched_gpu_s": bat_gpu, "sequential_s": seq_s,
            "kernel_s": kern_s, "kernel_bytes": cells * 19 * 8, "kernel_tag": list(plan.tag),
            "kernel_blocks": int(plan.plan.nblocks), "same": bool(same), "build": b, "cpu_start": cpu_start,
            "firsts": firsts, "timed_counts": timed_counts,
            "extra": {
                "instances": n_inst, "timed_physical_steps": steps, "warmup_physical_steps": warmup, "tol": tol,
                "k_smooth": k_smooth,
                "sizes": {str(k): v for k, v in sorted(collections.Counter(s["n"] for s in specs).items())},
                "levels": {str(k): v for k, v in sorted(collections.Counter(levels).items())},
                "total_fine_cells": cells,
                "batched_s": round(bat_s, 4), "sequential_s": round(seq_s, 4),
                "speedup_batched_over_sequential": round(seq_s / bat_s, 3),
                "instance_steps_per_s_batched": round(n_inst * steps / bat_s, 2),
                "instance_steps_per_s_sequential": round(n_inst * steps / seq_s, 2),
                "mlups_batched": round(cells * steps / bat_s / 1e6, 3),
                "mlups_sequential": round(cells * steps / seq_s / 1e6, 3),
                "fine_cell_vcycles_per_s_millions_batched": round(fine_cycles / bat_s / 1e6, 1),
                "fine_cell_vcycles_per_s_millions_sequential": round(fine_cycles / seq_s / 1e6, 1),
                "vcycles_timed_total": int(sum(counts)), "vcycles_max": int(max(counts)),
                "vcycles_median": int(sorted(counts)[len(counts) // 2]),
                "solves_reaching_tol": int(sum(1 for r in rels if r <= tol)), "solves": len(rels),
                "bitwise_equal_to_sequential": bool(same), "depth_tuning_s": round(b["tune_s"], 2),
                "batched_host_split_s": {k: round(v, 4) for k, v in bat.timing.items()},
                "batch_smooth_rows": {str(k): v for k, v in bat.batch.smooth_rows.items()},
                "per_instance": [[s_["n"], L, [int(c) for c in cs]] for s_, L, cs in zip(specs, levels, timed_counts)]}}


def c5_cpu_baseline(torch, r, seconds_cap=30.0, threads=None):
    """The CPU restatement timed on a bounded sample of the timed workload.  For the
    first instance of every (size, depth) pair of the set, the oracle runs the first
    timed physical step from the GPU's state twice, with 1 and with 3 V-cycles; the
    difference gives its seconds per V-cycle, the rest the step's glue.  The seconds
    the CPU would need for the timed workload are then those per-pair costs times the
    V-cycles every instance needed in every timed step (the GPU solves' counts: the
    same work), and MLUPS is fine cells x physical steps over that."""
    import problems as P
    b = r["build"]
    specs, levels, fields = b["specs"], b["levels"], b["fields"]
    warmup = r["extra"]["warmup_physical_steps"]
    threads = threads or os.cpu_count() or 1
    torch.set_num_threads(threads)
    cost = {}
    spent = 0.0
    for (n, L), i in sorted(r["firsts"].items()):
        fe, fd, dde = (x.permute(1, 2, 0).contiguous() for x in r["cpu_start"][i])
        o = P.biot_c5("oracle", specs[i], L, None, 10, K=fields[i], k_smooth=b["k_smooth"])
        o.elasticityTol = None
        ts = {}
        for cyc in (1, 1, 3):  # the first 1-cycle step warms the oracle up
            o.elasticitySteps = cyc
            t0 = time.perf_counter()
            o.step(fe, fd, dde, warmup)
            ts[cyc] = time.perf_counter() - t0
            spent += ts[cyc]
        per_cycle = max((ts[3] - ts[1]) / 2, 0.0)
        cost[n, L] = (max(ts[1] - per_cycle, 0.0), per_cycle)
        if spent > seconds_cap:
            break
    if len(cost) < len(r["firsts"]):
        return None
    total = 0.0
    for s, L, cs in zip(specs, levels, r["timed_counts"]):
        glue, per_cycle = cost[s["n"], L]
        total += sum(glue + c * per_cycle for c in cs)
    steps = r["extra"]["timed_physical_steps"]
    return dict(value=r["cells"] * steps / total / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=(f"oracle (torch CPU, {threads} threads, float64): one physical step with 1 and with 3 "
                        f"V-cycles for the first instance of each of the {len(cost)} (size, depth) pairs of the set "
                        f"({spent:.1f} s); the timed workload ({steps} physical steps of 64 instances, "
                        f"{sum(sum(c) for c in r['timed_counts'])} V-cycles) priced at those per-step and "
                        f"per-V-cycle costs: {total:.1f} CPU s"))


def _eta(f):
    """Displacement moments (m10, m01) of SoA D2Q8 populations, differentiable."""
    import torch
    e0 = f[0] - f[2] + f[4] - f[5] - f[6] + f[7]
    e1 = f[1] - f[3] + f[4] + f[5] - f[6] - f[7]
    return torch.stack((e0, e1))


def bench_adjoint(torch, n=1024, steps=1000, K=50, precision="f32/f32"):
    """Gradient cost against forward cost on the 2-D consolidation geometry.

    Elasticity: (n) x (n + 2) cells, fixed bottom, traction top, periodic in x, per-cell
    lambda = 2 and mu = 1 fields; misfit 1/2 |eta - eta*|^2 of the displacement after
    ``steps`` pseudo-steps, eta* from the same solve with mu perturbed by a seeded
    log-normal factor (seed 2).  Darcy: n x n (+ ghost layers), log-normal K (variance
    2, correlation 32, seed 1), omega = 1/(3K + 1/2); misfit 1/2 |p - p*|^2, p* from
    K perturbed (seed 2).  Gradients with respect to mu(x) and K(x) through the
    checkpointed adjoint (checkpoints every K steps, graph-replayed segments); the
    state is FP32, the gradient accumulators FP64."""
    import problems as P
    from src.adjoint import darcy_solve, elasticity_solve, ElasticityStepper, DarcyStepper
    out = {}
    S = 8 if precision == "f64/f64" else 4
    dt = torch.float64 if S == 8 else torch.float32

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    # ---- elasticity
    el = P.consolidation_elasticity("gpu", nx=n, precision=precision)
    nx, ny = el.nx, el.ny
    lam = torch.full((nx, ny), 2.0, dtype=dt, device="cuda")
    mu_t = torch.full((nx, ny), 1.0, dtype=dt, device="cuda")
    pert = torch.as_tensor(P.lognormal_field(nx, ny, 0.1, 32.0, 2), dtype=dt, device="cuda")
    f0 = torch.zeros((8, nx, ny), dtype=dt, device="cuda")
    with torch.no_grad():
        target = _eta(elasticity_solve(el, lam, mu_t * pert, f0, steps, checkpoint_every=K))
    mu = mu_t.clone().requires_grad_(True)
    # warm-up forward + backward: the solve's buffers and graphs are captured once
    (0.5 * ((_eta(elasticity_solve(el, lam, mu, f0, steps, checkpoint_every=K)) - target) ** 2).sum()).backward()
    mu.grad = None
    fN, t_fwd = timed(lambda: elasticity_solve(el, lam, mu, f0, steps, checkpoint_every=K))
    J = 0.5 * ((_eta(fN) - target) ** 2).sum()
    _, t_bwd = timed(lambda: J.backward())
    # plain forward (no checkpoints) for reference
    _, t_plain = timed(lambda: el.run_steps(f0.permute(1, 2, 0), steps, chunk=K))
    # adjoint kernel alone
    st = ElasticityStepper(el, lam.detach(), mu.detach())
    a0, a1 = torch.randn_like(f0), torch.empty_like(f0)
    g = st.grad_buffers(False)
    kern = event_pair_time(torch, lambda: st.adjoint(a0, f0, a1, g), 50)
    adj_bytes = nx * ny * (26 * S + 32)
    out["elasticity"] = {"grid": [nx, ny], "steps": steps, "checkpoint_every": K, "precision": precision,
                         "forward_s": round(t_fwd, 4), "plain_forward_s": round(t_plain, 4),
                         "backward_s": round(t_bwd, 4), "backward_over_forward": round(t_bwd / t_fwd, 3),
                         "gradient_over_forward": round((t_fwd + t_bwd) / t_fwd, 3),
                         "adjoint_kernel_us": round(kern * 1e6, 2),
                         "adjoint_kernel_gbs": round(adj_bytes / kern / 1e9, 1),
                         "adjoint_bytes_per_cell": 26 * S + 32,
                         "grad_norm": float(mu.grad.double().norm()), "misfit": float(J.detach())}
    log(f"  elasticity adjoint: forward {t_fwd:.3f} s, gradient {t_bwd:.3f} s")

    # ---- Darcy
    d = P.darcy_c2("gpu", n=n, precision=precision, heterogeneous=False)
    nxd, nyd = d.nx, d.ny
    Kf = torch.as_tensor(P.lognormal_field(nxd, nyd, 2.0, 32.0, 1), dtype=dt, device="cuda")
    Kp = Kf * torch.as_tensor(P.lognormal_field(nxd, nyd, 0.1, 32.0, 2), dtype=dt, device="cuda")
    fd0 = torch.as_tensor(d.assign_fields_sharded()).to(dt).permute(2, 0, 1).contiguous()
    with torch.no_grad():
        ptarget = darcy_solve(d, 1.0 / (3.0 * Kp + 0.5), fd0, steps, checkpoint_every=K).sum(0)
    Kg = Kf.clone().requires_grad_(True)
    (0.5 * ((darcy_solve(d, 1.0 / (3.0 * Kg + 0.5), fd0, steps, checkpoint_every=K).sum(0) - ptarget) ** 2)
     .sum()).backward()
    Kg.grad = None
    fdN, t_fwd = timed(lambda: darcy_solve(d, 1.0 / (3.0 * Kg + 0.5), fd0, steps, checkpoint_every=K))
    Jd = 0.5 * ((fdN.sum(0) - ptarget) ** 2).sum()
    _, t_bwd = timed(lambda: Jd.backward())
    std = DarcyStepper(d, (1.0 / (3.0 * Kf + 0.5)).detach())
    b0, b1 = torch.randn_like(fd0), torch.empty_like(fd0)
    gd = std.grad_buffers(False)
    kern = event_pair_time(torch, lambda: std.adjoint(b0, fd0, b1, gd), 50)
    adj_bytes = nxd * nyd * (28 * S + 16)
    out["darcy"] = {"grid": [nxd, nyd], "steps": steps, "checkpoint_every": K, "precision": precision,
                    "forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4),
                    "backward_over_forward": round(t_bwd / t_fwd, 3),
                    "gradient_over_forward": round((t_fwd + t_bwd) / t_fwd, 3),
                    "adjoint_kernel_us": round(kern * 1e6, 2), "adjoint_kernel_gbs": round(adj_bytes / kern / 1e9, 1),
                    "adjoint_bytes_per_cell": 28 * S + 16,
                    "grad_norm": float(Kg.grad.double().norm()), "misfit": float(Jd.detach())}
    log(f"  darcy adjoint: forward {t_fwd:.3f} s, gradient {t_bwd:.3f} s")
    return out


def bench_biot_adjoint(torch, n=1024, steps=50, ne=2, levels=None, K=50, precision="f32/f32", k_smooth=8):
    """Gradient of a displacement / pressure misfit through the coupled multigrid
    consolidation solve (2-D consolidation: n x (n + 2) cells, fixed undrained bottom,
    traction on the drained top; multigrid depth by the rule of
    examples/poroelasticity/2D.py:216-219 (10 levels for n = 1024), ``ne`` V-cycles
    per physical step as there).  ``k_smooth`` 8: with lambda = 2, mu = 1 the V-cycle
    with the driver's 4 smoothing steps diverges from 3 levels on (residual x 1e3 in
    100 cycles at 1024, 4 levels, FP32), with 8 it converges at every depth.
    Log-normal K (variance 2, correlation 32, seed 1),
    lambda = 2 and mu = 1 held as fields; target (eta*, p*) from the forward solve with
    K perturbed by a log-normal factor (variance 0.1, seed 2).  ``steps`` physical
    steps, checkpoints every ``K`` V-cycles; state FP32, accumulators FP64."""
    import problems as P
    from src.biot_adjoint import biot_solve
    dt = torch.float64 if precision == "f64/f64" else torch.float32
    if levels is None:
        # examples/poroelasticity/2D.py:216-219
        levels = int(n).bit_length() - 1
        while n % 2 ** (levels - 1):
            levels -= 1
    prob = P.biot_consolidation("gpu", nx=n, levels=levels, steps=ne, k_smooth=k_smooth, precision=precision)
    nx, ny = prob.elasticity.nx, prob.elasticity.ny
    Kf = 0.5 * torch.as_tensor(P.lognormal_field(nx, ny, 2.0, 32.0, 1), dtype=torch.float64)
    Kt = Kf * torch.as_tensor(P.lognormal_field(nx, ny, 0.1, 32.0, 2), dtype=torch.float64)
    lam = torch.full((nx, ny), 2.0, dtype=dt, device="cuda")
    mu = torch.full((nx, ny), 1.0, dtype=dt, device="cuda")

    def omega(k):
        return (1.0 / (3.0 * k + 0.5)).to(device="cuda", dtype=dt)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    with torch.no_grad():
        te, tp = biot_solve(prob, omega(Kt), lam, mu, steps, checkpoint_every=K)
        te, tp = te.clone(), tp.clone()
    Kg = Kf.to(device="cuda", dtype=dt).requires_grad_(True)
    mug = mu.clone().requires_grad_(True)

    def misfit():
        e, q = biot_solve(prob, 1.0 / (3.0 * Kg + 0.5), lam, mug, steps, checkpoint_every=K)
        return 0.5 * ((e - te) ** 2).sum(dtype=torch.float64) + 0.5 * ((q - tp) ** 2).sum(dtype=torch.float64)
    misfit().backward()  # warm-up: session buffers and graphs
    Kg.grad = mug.grad = None
    torch.cuda.reset_peak_memory_stats()
    J, t_fwd = timed(misfit)
    _, t_bwd = timed(lambda: J.backward())
    finite = bool(torch.isfinite(Kg.grad).all() and torch.isfinite(mug.grad).all())
    s = next(iter(prob._biot_adjoint_sessions.values()))
    return {"grid": [nx, ny], "levels": levels, "k_smooth": k_smooth, "vcycles_per_step": ne, "physical_steps": steps,
            "checkpoint_every_vcycles": K, "checkpointed_steps": len(s.ckpt_step), "precision": precision,
            "forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4),
            "backward_over_forward": round(t_bwd / t_fwd, 3),
            "gradient_over_forward": round((t_fwd + t_bwd) / t_fwd, 3),
            "peak_memory_gb": round(torch.cuda.max_memory_allocated() / 1e9, 2),
            "misfit": float(J.detach()), "grad_K_norm": float(Kg.grad.double().norm()),
            "grad_mu_norm": float(mug.grad.double().norm()), "finite": finite}


def bench_steady_adjoint(torch, n=512, tol=1e-8, check_every=1000, checks=10):
    """Adjoint / forward time at a relative residual of 1e-8 on C2 (Darcy 514^2, FP64,
    log-normal K of variance 1 and correlation 16, seed 0).  Forward: steps to the
    steady state (``run_steps_to``: the tuned tile launches, the residual tested on the
    device).  Gradient of J = 1/2 |p - p*|^2 with respect to the relaxation rate
    omega(x) (so to K) by the fixed-point adjoint ``a = (dS/df)^T a + dJ/df`` iterated
    to the same relative residual (``darcy_steady_adjoint``: one adjoint step per
    launch, device-side test), then one parameter pass; p* from the steady state with
    K perturbed by a log-normal factor (variance 0.1, seed 2).  Both loops test every
    ``check_every`` steps and read the device-side latch every ``checks`` tests, as
    ``c2_time_to_solution``; each solve runs once untimed first (graph captures).
    Both residuals are relative to the first step's: the adjoint's first step is
    ``(dS/df)^T dJ/df``, a smooth field, so it needs more steps than the forward (whose
    first step is the start's transient) to fall by the same factor."""
    import problems as P
    from src.adjoint import darcy_steady_adjoint

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    d = P.darcy_c2("gpu", n=n, precision="f64/f64", heterogeneous=True, seed=0, variance=1.0, corr=16.0)
    d.autotune()
    nx, ny = d.nx, d.ny
    om = d.omega_field.clone()
    K = (1.0 / om - 0.5) / 3.0
    Kp = K * torch.as_tensor(P.lognormal_field(nx, ny, 0.1, 16.0, 2), dtype=K.dtype, device=K.device).reshape(nx, ny, 1)
    f0 = d.assign_fields_sharded()
    mx = 4000000
    d.set_omega_field(1.0 / (3.0 * Kp + 0.5))
    ft, kt, _ = d.run_steps_to(f0, tol, check_every, max_steps=mx, checks=checks)
    p_star = ft.sum(-1)
    d.set_omega_field(om)
    d.run_steps_to(f0, tol, check_every, max_steps=check_every * checks + 1, checks=checks)
    (fs, kf, rf), t_fwd = timed(lambda: d.run_steps_to(f0, tol, check_every, max_steps=mx, checks=checks))
    dJdf = (fs.sum(-1) - p_star).unsqueeze(0).expand(9, nx, ny).contiguous()
    fs_soa = fs.permute(2, 0, 1).contiguous()
    om2 = om.reshape(nx, ny)
    darcy_steady_adjoint(d, om2, fs_soa, dJdf, tol=tol, max_steps=check_every * checks + 1, check_every=check_every)
    (res, t_adj) = timed(lambda: darcy_steady_adjoint(d, om2, fs_soa, dJdf, tol=tol, max_steps=mx,
                                                      check_every=check_every))
    grads, _, ka, ra = res
    gK = grads["omega"].reshape(nx, ny) * (-3.0 * om.reshape(nx, ny).double() ** 2)
    return {"grid": [nx, ny], "tol": tol, "precision": "f64/f64",
            "forward_steps": kf, "forward_s": round(t_fwd, 4), "forward_rel_residual": rf,
            "adjoint_steps": ka, "adjoint_s": round(t_adj, 4), "adjoint_rel_residual": ra,
            "adjoint_over_forward": round(t_adj / t_fwd, 3),
            "gradient_over_forward": round((t_fwd + t_adj) / t_fwd, 3),
            "target_steps": kt, "check_every": check_every, "host_reads_every_steps": check_every * checks,
            "grad_K_norm": float(gK.norm()),
            "misfit": float(0.5 * ((fs.sum(-1) - p_star).double() ** 2).sum())}


def profiled_traffic(kernel, blocks, cells):
    """L2-to-memory bytes per launch of this kernel configuration from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, calibrated on a copy), or None
    when that exact configuration was not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    rec = table.get(f"{kernel}|blocks={blocks}|cells={cells}")
    return None if rec is None else rec["read_bytes"] + rec["write_bytes"]


def c2_cpu_record(n, precision, seconds, threads=None):
    """The CPU restatement's C2 rate (Darcy steps on the CPU; all host cores unless
    ``threads`` is given)."""
    import torch
    import problems as P
    threads = threads or os.cpu_count() or 1
    torch.set_num_threads(threads)
    o = P.darcy_c2("oracle", n=n, precision=precision, heterogeneous=True, seed=0, variance=1.0, corr=16.0)
    f = o.assign_fields_sharded()
    f, _ = o.step(f, 0)  # warm-up (the reference starts its clock after step 1)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds:
        f, _ = o.step(f, k + 1)
        k += 1
    el = time.perf_counter() - t0
    return dict(value=o.nx * o.ny * k / el / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"{k} Darcy steps of the {o.nx}x{o.ny} FP64 workload on the CPU restatement "
                       f"(torch CPU, {threads} threads), {el:.1f} s")


def c1_cpu_record(threads=None, tol=1e-10, check_every=10, max_steps=20000):
    """C1 on the CPU restatement, the whole solve: plain steps of the periodic 64^2
    elasticity from the zero state until |f_k - f_(k-1)| < tol |f_1 - f_0|, tested
    every ``check_every`` steps (k = 1 + j check_every), the rule of the GPU run."""
    import torch
    import problems as P
    threads = threads or os.cpu_count() or 1
    torch.set_num_threads(threads)
    o = P.periodic_elasticity("oracle", nx=64)
    f0 = o.assign_fields_sharded()
    o.step(f0, 0)  # warm-up (the reference starts its clock after step 1)
    t0 = time.perf_counter()
    f, _ = o.step(f0, 0)
    r1 = float(torch.linalg.norm((f - f0).reshape(-1)))
    k, ratio = 1, 1.0
    while k < max_steps:
        for _ in range(check_every - 1):
            f, _ = o.step(f, k)
            k += 1
        prev = f
        f, _ = o.step(f, k)
        k += 1
        ratio = float(torch.linalg.norm((f - prev).reshape(-1))) / r1
        if ratio < tol:
            break
    el = time.perf_counter() - t0
    return dict(value=64 * 64 * k / el / 1e6, unit="MLUPS", cores=threads, kind="port", steps=k,
                seconds=round(el, 3), rel_residual=ratio,
                sample=f"the whole C1 solve ({k} steps of the 64x64 FP64 workload to {tol:g}) on the CPU restatement "
                       f"(torch CPU, {threads} threads), {el:.2f} s")


def _oracle_step_costs(o, state, t, cycles=(1, 1, 3)):
    """Seconds of one physical step of an oracle Biot problem with 1 and with 3 inner
    V-cycles (the first 1-cycle step warms the oracle up): (glue, per V-cycle, spent)."""
    ts, spent = {}, 0.0
    o.elasticityTol = None
    for cyc in cycles:
        o.elasticitySteps = cyc
        t0 = time.perf_counter()
        o.step(*state, t)
        ts[cyc] = time.perf_counter() - t0
        spent += ts[cyc]
    per_cycle = max((ts[3] - ts[1]) / 2, 0.0)
    return max(ts[1] - per_cycle, 0.0), per_cycle, spent


def c3_cpu_record(gpu, threads=None, n=256, levels=8):
    """The CPU restatement on the periodic Biot workload the GPU runs with the reference
    driver's material (``gpu``: that run's record): one physical step with 1 and with 3
    V-cycles gives the step's glue and the cost per V-cycle; the GPU run's physical
    steps and V-cycles are priced at those costs."""
    import torch
    import problems as P
    threads = threads or os.cpu_count() or 1
    torch.set_num_threads(threads)
    lam, mu = 0.11 * 0.8 / (1 - 0.8 ** 2), 0.5 * 0.11 / 1.8
    o = P.biot_periodic("oracle", nx=n, levels=levels, steps=1, lam=lam, mu=mu, K=0.1)
    glue, per_cycle, spent = _oracle_step_costs(o, o.initial_state(), 1)
    steps, cycles = gpu["physical_steps"], gpu["vcycles"]
    total = steps * glue + cycles * per_cycle
    return dict(value=n * n * steps / total / 1e6, unit="MLUPS", cores=threads, kind="port",
                seconds_priced=round(total, 2), per_vcycle_s=round(per_cycle, 4), step_glue_s=round(glue, 4),
                sample=(f"oracle (torch CPU, {threads} threads, float64): one physical step of the {n}x{n} Biot "
                        f"workload ({levels} levels) with 1 and with 3 V-cycles ({spent:.1f} s); the GPU run's "
                        f"{steps} physical steps and {cycles} V-cycles priced at those costs: {total:.1f} CPU s"))


def c4_cpu_record(gpu, threads=None, n=1024, k_smooth=8):
    """The CPU restatement on the C4 forward solve (2-D consolidation, 1024 x 1026,
    FP32, log-normal K, lambda and mu as fields): one physical step with 1 and with 3
    V-cycles; the GPU run's forward (its physical steps, ``vcycles_per_step`` each) is
    priced at those costs.  The gradient is not timed on the CPU."""
    import torch
    import problems as P
    threads = threads or os.cpu_count() or 1
    torch.set_num_threads(threads)
    levels, steps, ne = gpu["levels"], gpu["physical_steps"], gpu["vcycles_per_step"]
    nx, ny = n, n + 2
    K = 0.5 * torch.as_tensor(P.lognormal_field(nx, ny, 2.0, 32.0, 1), dtype=torch.float64)
    o = P.biot_consolidation("oracle", nx=n, levels=levels, steps=ne, k_smooth=k_smooth, precision="f32/f32",
                             lambda_field=torch.full((nx, ny), 2.0, dtype=torch.float32),
                             mu_field=torch.full((nx, ny), 1.0, dtype=torch.float32),
                             omega_field=(1.0 / (3.0 * K + 0.5)).to(torch.float32)[..., None])
    glue, per_cycle, spent = _oracle_step_costs(o, o.initial_state(), 0)
    total = steps * glue + steps * ne * per_cycle
    return dict(value=nx * ny * steps / total / 1e6, unit="MLUPS", cores=threads, kind="port",
                seconds_priced=round(total, 2), per_vcycle_s=round(per_cycle, 3), step_glue_s=round(glue, 3),
                gpu_forward_s=gpu.get("forward_s"),
                sample=(f"oracle (torch CPU, {threads} threads, float32 state): one physical step of the C4 "
                        f"consolidation ({nx}x{ny}, {levels} levels, k = {k_smooth}) with 1 and with 3 V-cycles "
                        f"({spent:.1f} s); the GPU forward's {steps} steps x {ne} V-cycles priced at those costs: "
                        f"{total:.1f} CPU s"))


def only(torch, args):
    """Counter-profile entry: one kernel family, ``--steps`` launches."""
    n = 4096
    if args.only == "c5_kernel":
        # the headline's dominant launch alone, every instance active (the batch's own
        # construction runs one eager cycle with every instance active first)
        b = c5_build(torch, n_inst=args.instances, sequential=False)
        kern_s, plan = c5_kernel_time(torch, b["bat"], args.steps)
        cells = sum(s["n"] ** 2 for s in b["specs"])
        log(f"c5 dominant launch {plan.tag} ({plan.plan.nblocks} blocks): {kern_s * 1e6:.1f} us, "
            f"{cells * 152 / kern_s / 1e9:.0f} GB/s compulsory")
    elif args.only == "copy":
        # plain device copy of a known size: 1 GiB read + 1 GiB written per launch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy: {args.steps} launches of {a.numel() * 8} B")
    elif args.only == "c2":
        r = bench_darcy(torch, args.n, "f64/f64", args.steps, chunk=48, tile_steps=args.tile_steps,
                        tile_threads=args.tile_threads)
        d = r["solver"]
        log(f"c2 {d.nx}x{d.ny}: {r['launch_kind']} launches of {r['steps_per_launch']} steps "
            f"({r['tile_threads']} threads per tile), "
            f"{r['kernel_s'] * 1e6:.2f} us/launch, {r['kernel_s'] * 1e6 / r['steps_per_launch']:.2f} us/step")
    elif args.only == "biot_vcycle":
        r = bench_biot_vcycle(torch, reps=args.steps)
        log(f"biot vcycle: {r}")
    elif args.only == "c1":
        print(json.dumps(bench_c1(torch)), flush=True)
    elif args.only == "biot_vcycle_eager":
        # the same V-cycle launched eagerly (no graph replay), for kernel traces
        r = bench_biot_vcycle(torch, reps=args.steps, eager=True)
        log(f"biot vcycle (eager launches): {r}")
    elif args.only == "c5":
        r = bench_c5(torch, steps=args.steps, warmup=args.warmup, n_inst=args.instances, kernel_reps=args.kernel_reps)
        print(json.dumps(dict(r["extra"], kernel_us=round(r["kernel_s"] * 1e6, 2))), flush=True)
    elif args.only == "adjoint":
        print(json.dumps(bench_adjoint(torch)), flush=True)
    elif args.only == "biot_adjoint":
        print(json.dumps(bench_biot_adjoint(torch, n=args.n, steps=args.steps)), flush=True)
    elif args.only == "steady_adjoint":
        print(json.dumps(bench_steady_adjoint(torch, n=args.n)), flush=True)
    elif args.only == "large_elast":
        r = bench_large_elasticity(torch, n, args.steps)
        log(f"elasticity {n}^2: {r['steps_per_launch']} steps per launch, {r['kernel_s'] * 1e6:.1f} us/launch, "
            f"{r['kernel_s'] * 1e6 / r['steps_per_launch']:.1f} us/step, {r['bytes'] / r['kernel_s'] / 1e9:.0f} GB/s")
    else:
        r = bench_darcy(torch, n - 2, "f64/f64", args.steps, chunk=16, tile_steps=args.tile_steps,
                        tile_threads=args.tile_threads)
        log(f"darcy {n}^2: {r['launch_kind']} launches of {r['steps_per_launch']} steps "
            f"({r['tile_threads']} threads per tile), "
            f"{r['kernel_s'] * 1e6:.1f} us/launch, {r['kernel_s'] * 1e6 / r['steps_per_launch']:.1f} us/step, "
            f"{r['bytes'] / r['kernel_s'] / 1e9:.0f} GB/s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed physical steps of the batched-Biot headline")
    ap.add_argument("--warmup", type=int, default=2, help="untimed physical steps before them")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--instances", type=int, default=64, help="instances of the batched Biot headline")
    ap.add_argument("--kernel-reps", type=int, default=20, help="launches timed for the dominant kernel")
    ap.add_argument("--tile-steps", type=int, default=0,
                    help="with --only c2 or large_darcy: steps per launch of the LDS-tile kernel instead of the tuned launch")
    ap.add_argument("--tile-threads", type=int, default=0, choices=[0, 256, 512, 704],
                    help="with --tile-steps: threads per tile block, which fixes the tile shape (256, 512: "
                         "32 x 32; 704: 44 x 48; 0: the library's choice)")
    ap.add_argument("--only", choices=["large_elast", "large_darcy", "copy", "c2", "c5", "c5_kernel", "adjoint",
                                       "biot_vcycle", "biot_vcycle_eager", "biot_adjoint", "c1", "steady_adjoint"],
                    help="run one kernel family only (for counter profiles); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path shards by x-slabs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(int(os.environ.get("LOCAL_RANK", "0")))
    if args.only:
        return only(torch, args)

    log(f"C5 batched Biot: {args.instances} instances, warmup {args.warmup}, steps {args.steps}")
    r = bench_c5(torch, steps=args.steps, warmup=args.warmup, n_inst=args.instances, kernel_reps=args.kernel_reps)
    cells = r["cells"]
    value = cells * args.steps / r["batched_s"] / 1e6
    achieved = r["kernel_bytes"] / r["kernel_s"] / 1e9
    traffic = profiled_traffic(C5_KERNEL, r["kernel_blocks"], cells)
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["batched_s"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: 64 seeded instances (grid sizes, Lame parameters, log-normal permeability fields; seed 3)",
        "config": {"workload": (f"C5 batched poroelasticity: {args.instances} periodic Biot instances, grids 64..512 "
                                "cells per side, FP64, V-cycles (8 smoothing steps) to 1e-8 of |S(0)| in every "
                                "physical step; batched vs one after another"),
                   "instances": args.instances, "total_fine_cells": cells, "precision": "f64/f64", "inner_tol": 1e-8,
                   "k_smooth": 8, "step": "one physical step of every instance",
                   "mlups_convention": "fine cells x physical steps / s / 1e6 (reference poroelasticity.py:288-299)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel and launch grid, "
                                        "calibrated on a copy (profiles/c5_batch_smooth_hbm_traffic.md)")
                     if traffic else "not profiled for this launch grid",
                     "kernel": C5_KERNEL, "launch": "level-0 smoothing of all instances, 2 steps per launch",
                     "launch_blocks": r["kernel_blocks"], "kernel_us": round(r["kernel_s"] * 1e6, 2),
                     "timed": f"{args.kernel_reps} launches with every instance active, HIP events on their stream",
                     "bytes_per_cell_per_launch": 152, "cells": cells, "algorithmic_bytes_per_launch": r["kernel_bytes"]},
    }
    if not r["same"]:
        log("WARNING: the batched and the sequential final states differ")
    extras = {"c5_batched_biot_fp64": r["extra"]}

    def extra(name, fn):
        # an extra that fails is reported, and the headline line is still printed
        try:
            extras[name] = fn()
        except Exception as e:  # noqa: BLE001
            log(f"extra {name} failed: {type(e).__name__}: {e}")
            extras[name] = {"error": f"{type(e).__name__}: {e}"}

    if not args.no_extras:
        from src import autotune
        log("C2 plain Darcy 512^2, FP64 and FP32")
        extra("c2_darcy_512", lambda: c2_record(torch, args.n))
        extra("c2_darcy_512_to_1e-8_fp64", lambda: c2_time_to_solution(torch, args.n))
        log("C1 64^2 to 1e-10")
        extra("c1_elasticity_64_fp64", lambda: bench_c1(torch))
        log("C3 Biot 256^2 MG (4 levels, lambda = mu = 1), inner 1e-8")
        extra("c3_biot_256_mg4_fp64", lambda: bench_c3(torch, steps=200, max_cycles=60))
        lam, mu = 0.11 * 0.8 / (1 - 0.8 ** 2), 0.5 * 0.11 / 1.8
        log("Biot 256^2 MG (8 levels, material of the reference periodic driver), inner 1e-8")
        extra("c3_biot_256_mg8_ref_material_fp64",
              lambda: bench_c3(torch, levels=8, steps=200, lam=lam, mu=mu, K=0.1))
        log("adjoint: consolidation geometry 1024, FP32, 1000 pseudo-steps, checkpoints every 50")
        extra("adjoint_1024_fp32", lambda: bench_adjoint(torch))
        log("C4: gradient through the coupled multigrid consolidation, 1024 x 1026, FP32, 50 steps")
        extra("c4_biot_adjoint_1024_fp32", lambda: bench_biot_adjoint(torch))
        log("device copy 1 GiB (achievable HBM bandwidth)")
        extra("stream_copy", lambda: bench_copy(torch))
        log("Biot 4096^2 fused-coupling V-cycle (6 levels, k = 2), fp64")
        extra("biot_vcycle_4096_fp64", lambda: bench_biot_vcycle(torch))
        for n in (4096,):
            log(f"elasticity {n}^2 fp64 (HBM-resident)")
            rl = bench_large_elasticity(torch, n, 8)
            gbs = rl["bytes"] / rl["kernel_s"] / 1e9
            extras[f"elasticity_{n}_fp64"] = {"mlups": round(rl["cells"] * rl["steps"] / rl["elapsed"] / 1e6, 1),
                                              "kernel_us": round(rl["kernel_s"] * 1e6, 2), "gbs": round(gbs, 1),
                                              "steps_per_launch": rl["steps_per_launch"],
                                              "threads_per_block": rl["tile_threads"],
                                              "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 144}
            log(f"darcy {n}^2 fp64 (HBM-resident)")
            rd = bench_darcy(torch, n - 2, "f64/f64", 8, chunk=32)
            gbs = rd["bytes"] / rd["kernel_s"] / 1e9
            extras[f"darcy_{n}_fp64"] = {"mlups": round(rd["cells"] * rd["steps"] / rd["elapsed"] / 1e6, 1),
                                         "kernel_us": round(rd["kernel_s"] * 1e6, 2), "gbs": round(gbs, 1),
                                         "steps_per_launch": rd["steps_per_launch"],
                                         "launch_kind": rd["launch_kind"], "threads_per_block": rd["tile_threads"],
                                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 152}
            if rd["launch_kind"] == "tile":
                tx_, ty_ = autotune.TILE_SHAPES[rd["tile_threads"]]
                kn = f"k_tile_steps<double, xb::DarcyColl<double>, 2, {tx_}, {ty_}, {rd['tile_threads']}>"
                extras[f"darcy_{n}_fp64"]["traffic"] = profiled_traffic(kn, rd["tile_blocks"], rd["cells"])
    if not args.no_cpu:
        log("cpu baseline: the oracle on a sample of the batched-Biot workload")
        line["cpu_baseline"] = c5_cpu_baseline(torch, r)
        if not args.no_extras:
            log(f"cpu: the oracle on C2 ({args.cpu_seconds:.0f} s)")
            extra("c2_cpu_oracle", lambda: c2_cpu_record(args.n, "f64/f64", args.cpu_seconds))
            # one thread: the CPU restatement without the host's parallelism
            extra("c2_cpu_oracle_1_thread", lambda: c2_cpu_record(args.n, "f64/f64", args.cpu_seconds / 2, threads=1))
            extra("c5_cpu_oracle_1_thread", lambda: c5_cpu_baseline(torch, r, seconds_cap=40.0, threads=1))
            log("cpu: the oracle on C1 (the whole solve), C3 and the C4 forward (priced samples)")
            extra("c1_cpu_oracle", lambda: c1_cpu_record())
            extra("c1_cpu_oracle_1_thread", lambda: c1_cpu_record(threads=1))
            c3 = extras.get("c3_biot_256_mg8_ref_material_fp64", {})
            if "vcycles" in c3 and "diverged" not in c3:
                extra("c3_cpu_oracle", lambda: c3_cpu_record(c3))
            c4 = extras.get("c4_biot_adjoint_1024_fp32", {})
            if "levels" in c4:
                extra("c4_cpu_oracle_forward", lambda: c4_cpu_record(c4))
            torch.set_num_threads(os.cpu_count() or 1)
    line["extra"] = extras
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
