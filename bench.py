#!/usr/bin/env python
"""Benchmark of the lattice-Boltzmann Darcy / elasticity path on one MI355X.

Workload (the headline line): the plain D2Q9 Darcy solver on a 512 x 512 domain
(plus one ghost layer per wall: 514 x 514 cells), FP64, log-normal permeability
K = exp(g) with g a seeded Gaussian random field of variance 1 and exponential
covariance of correlation length 16 cells, p = 1 on the left and p = 0 on the
right (anti-bounce-back), no flux at the bottom and top (bounce-back).  One
"step" is one Darcy LBM update of the whole grid.  ``value`` is MLUPS in the
reference convention (cells x steps / s / 1e6), with the populations resident in
HBM; the steps replay captured HIP graphs of 64 updates, each launch advancing two
steps (``xb_darcy_step2``: populations read once, collided, streamed and
wall-corrected twice in registers, written once).

``roofline`` prices the dominant kernel (the two-step Darcy sweep) at its
algorithmic bytes per launch, 19 values per cell (9 populations read, 9 written,
the per-cell relaxation rate read) x 8 B, over its average launch duration measured
with HIP events on the launch stream; ``per_step_equivalent_gbs`` is what single
steps would have to sustain for the same step rate.  ``cpu_baseline`` times the CPU restatement of the
reference algorithm (torch CPU, float64, all host cores) on a bounded sample of
the same workload.
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


def timed_steps(torch, run, f, steps):
    """Run ``steps`` updates between synchronisations; returns (f, wall s, GPU s from
    HIP events recorded on the launch stream around the same region)."""
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    f = run(f, steps)
    e1.record()
    torch.cuda.synchronize()
    return f, time.perf_counter() - t0, e0.elapsed_time(e1) / 1e3


def bench_darcy(torch, n, precision, steps, warmup, chunk=64, tune=True, tile_steps=0, tile_threads=0):
    d = darcy_workload(n, precision)
    if tune:
        d.autotune()
    if tile_steps:
        from src import autotune
        d.step_launch = autotune.tile_choice(tile_steps, tile_threads)
    f = d.assign_fields_sharded()
    f = d.run_steps(f, warmup, chunk)
    f, wall, gpu = timed_steps(torch, lambda f, k: d.run_steps(f, k, chunk), f, steps)
    S = 8 if precision == "f64/f64" else 4
    cells = d.nx * d.ny
    # the timed graphs hold only launches of the tuned kind (m steps each: one or two
    # per sweep, or m on LDS tiles): GPU time / (steps / m) = average launch
    # duration; a launch reads and writes every population once (compulsory 19
    # values per cell)
    from src import autotune
    choice, spl, _ = d._multi()
    kind = "tile" if choice > autotune.TILE_BASE else ("sweep2" if choice == autotune.SWEEP2 else "sweep")
    threads, blocks = None, None
    if kind == "tile":
        threads = autotune.tile_threads(choice) or autotune.library_tile_threads(d.nx, d.ny, spl)
        blocks = autotune.tile_blocks(d.nx, d.ny, spl, threads)
    return dict(elapsed=wall, gpu=gpu, cells=cells, kernel_s=gpu / steps * spl, bytes=cells * 19 * S,
                steps_per_launch=spl, launch_kind=kind, tile_threads=threads, tile_blocks=blocks, solver=d)


def bench_large_elasticity(torch, n, steps, warmup, precision="f64/f64", chunk=32):
    """Periodic D2Q8 step on an n x n grid with a body force (18 values per cell)."""
    import problems as P
    el = P.periodic_elasticity("gpu", nx=n, precision=precision)
    el.autotune()
    f = el.assign_fields_sharded()
    f = el.run_steps(f, warmup, chunk)
    f, wall, gpu = timed_steps(torch, lambda f, k: el.run_steps(f, k, chunk), f, steps)
    S = 8 if precision == "f64/f64" else 4
    cells = n * n
    choice, spl, _ = el._multi(None)
    from src import autotune
    threads = None
    if choice > autotune.TILE_BASE:
        threads = autotune.tile_threads(choice) or autotune.library_tile_threads(n, n, spl)
    return dict(elapsed=wall, gpu=gpu, cells=cells, kernel_s=gpu / steps * spl, bytes=cells * 18 * S,
                steps_per_launch=spl, tile_threads=threads)


def residual_solve(torch, d, tol, max_steps, check_every=1000, chunk=64):
    """Steps until ||f_{k+1}-f_k|| / ||f_1-f_0|| < tol; returns (steps, seconds)."""
    f = d.assign_fields_sharded()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f1 = d.run_steps(f, 1)
    r0 = torch.linalg.norm(f1 - f).item()
    f = f1
    k = 1
    while k < max_steps:
        f = d.run_steps(f, check_every - 1, chunk)
        fn = d.run_steps(f, 1)
        r = torch.linalg.norm(fn - f).item()
        f = fn
        k += check_every
        if r / r0 < tol:
            break
    torch.cuda.synchronize()
    return k, time.perf_counter() - t0, r / r0


def bench_c1(torch, tol=1e-10, check_every=10, max_steps=20000):
    """C1: periodic 64^2 elasticity, FP64, plain steps until the relative displacement
    residual ||f_{k+1}-f_k|| / ||f_1-f_0|| < tol; error to the closed form."""
    import numpy as np
    import problems as P
    el = P.periodic_elasticity("gpu", nx=64)
    f0 = el.assign_fields_sharded()
    chunk = check_every - 1 if check_every % 2 else check_every - 2
    # the step runners (buffers and captured graphs) are built before the clock starts
    el.run_steps(f0, 1)
    el.run_steps(f0, check_every - 1, chunk=chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f1 = el.run_steps(f0, 1)
    r0 = torch.linalg.norm(f1 - f0).item()
    f, k, r = f1, 1, r0
    while k < max_steps:
        f = el.run_steps(f, check_every - 1, chunk=chunk)
        fn = el.run_steps(f, 1)
        r = torch.linalg.norm(fn - f).item()
        f, k = fn, k + check_every
        if r / r0 < tol:
            break
    torch.cuda.synchronize()
    el_s = time.perf_counter() - t0
    eta, _ = el.update_macroscopic(f)
    ue = P.manufactured_c1(64) / (8 * np.pi ** 2)
    err = float(np.linalg.norm(eta.cpu().numpy() - ue) / np.linalg.norm(ue))
    return {"steps": k, "seconds": round(el_s, 4), "mlups": round(64 * 64 * k / el_s / 1e6, 1),
            "rel_residual": r / r0, "rel_error_closed_form": round(err, 6)}


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
    for t in range(1, steps + 1):
        out = b._step_soa(fe, fd, dde, t)
        fe, fd, dde = out[0], out[1], out[4]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cycles = sum(b.inner_counts)
    rels = list(b.inner_residuals)
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
    return {"physical_steps": steps, "seconds": round(el, 4), "vcycles": cycles,
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
    restriction (11 S + 8 S / 4; coarse levels 17 S + 8 S / 4); coarse states are
    zero-filled (9 S) and prolonged (8 S / 4 read coarse, 16 S + 1 S per fine cell)."""
    total = 0.0
    launches = 2 * (-(-k // m))
    for lvl in range(levels):
        c = (n >> lvl) ** 2
        coarsest = lvl == levels - 1
        if m == 1:
            smooth = 20 if lvl == 0 else 26
        else:
            smooth = 19 if lvl == 0 else 25
        total += launches * smooth * c
        if not coarsest:
            total += ((11 if lvl == 0 else 17) + 2) * c   # residual + restriction out
            total += 9 * c / 4                             # zero fill of the coarse state
            total += (2 + 17) * c                          # prolongation
    return total * S


def bench_c5(torch, n_inst=64, steps=2, tol=1e-8, seed=3, max_cycles=3000, k_smooth=8):
    """Batched Biot: ``n_inst`` periodic Biot instances (sizes from {64, 128, 256, 512},
    lambda, mu ~ U[0.5, 2], log-permeability mean ~ U[-1, 1], log-normal K), FP64,
    ``steps`` physical steps each, elasticity V-cycles to the inner relative residual
    ``tol``.  The same instances run one after another (each with its captured
    V-cycle graph and a host-side residual test per cycle) and as one batch
    (batched launches, device-side convergence masks, captured batched cycles).
    Each instance's hierarchy depth is tuned first (shortest predicted time to the
    tolerance) and used by both runs.  The V-cycles smooth ``k_smooth`` = 8 times
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
    seq = [P.biot_c5("gpu", s, L, tol, max_cycles, K=K, k_smooth=k_smooth) for s, L, K in zip(specs, levels, fields)]
    for p in seq:
        p.prepare_inner()
    t0 = time.perf_counter()
    bat = BiotBatch([P.biot_c5("gpu", s, L, tol, max_cycles, K=K, k_smooth=k_smooth)
                     for s, L, K in zip(specs, levels, fields)])
    log(f"  batch built in {time.perf_counter() - t0:.1f} s, smoothing launch shapes {bat.batch.smooth_rows}")
    seq_states = [p.initial_state() for p in seq]
    bat_states = bat.initial_state()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq_out = []
    for p, st in zip(seq, seq_states):
        for t in range(steps):
            o = p._step_soa(*st, t)
            st = (o[0], o[1], o[4])
        seq_out.append(st)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    log(f"  sequential {seq_s:.2f} s")
    t0 = time.perf_counter()
    for t in range(steps):
        outs = bat.step(bat_states, t)
        bat_states = [(o[0], o[1], o[4]) for o in outs]
    torch.cuda.synchronize()
    bat_s = time.perf_counter() - t0
    log(f"  batched {bat_s:.2f} s")
    same = all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(seq_out, bat_states))
    # one batched cycle with every instance active, and one 512^2 instance's own cycle
    eb = bat.batch
    eb._reset()
    eb.rref.fill_(-1.0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eb.graphs[0].replay()
    e0.record()
    for r in range(6):
        eb.graphs[(r + 1) % 2].replay()
    e1.record()
    torch.cuda.synchronize()
    all_active = int(eb.nactive.item())
    batch_cycle_s = e0.elapsed_time(e1) / 1e3 / 6
    m_b = eb._tile_steps(0)
    batch_bytes = sum(vcycle_bytes(s_["n"], L, k_smooth, m=m_b) for s_, L in zip(specs, levels))
    big = max(range(n_inst), key=lambda i: (specs[i]["n"], -levels[i]))
    vg = seq[big]._inner
    vg.replay()
    e0.record()
    for _ in range(6):
        vg.replay()
    e1.record()
    torch.cuda.synchronize()
    one_cycle_s = e0.elapsed_time(e1) / 1e3 / 6
    one_bytes = vcycle_bytes(specs[big]["n"], levels[big], k_smooth,
                             m=seq[big].elasticity._smooth_tile_steps(0, seq[big].coupling_force()))
    per_step = [[p.inner_counts[t] for p in seq] for t in range(steps)]
    counts = [c for p in seq for c in p.inner_counts]
    rels = [r for p in seq for r in p.inner_residuals]
    fine_cycles = sum(s["n"] ** 2 * sum(p.inner_counts) for s, p in zip(specs, seq))
    return {"instances": n_inst, "physical_steps_each": steps, "tol": tol, "k_smooth": k_smooth,
            "sizes": {str(k): v for k, v in sorted(collections.Counter(s["n"] for s in specs).items())},
            "levels": {str(k): v for k, v in sorted(collections.Counter(levels).items())},
            "sequential_s": round(seq_s, 4), "batched_s": round(bat_s, 4), "speedup": round(seq_s / bat_s, 3),
            "instances_per_s_sequential": round(n_inst / seq_s, 3),
            "instances_per_s_batched": round(n_inst / bat_s, 3),
            "vcycles_total": int(sum(counts)), "vcycles_max": int(max(counts)),
            "vcycles_median": int(sorted(counts)[len(counts) // 2]),
            "vcycles_per_physical_step": [int(sum(c)) for c in per_step],
            "solves_reaching_tol": int(sum(1 for r in rels if r <= tol)), "solves": len(rels),
            "batched_fine_cell_vcycles_per_s_millions": round(fine_cycles / bat_s / 1e6, 1),
            "bitwise_equal_to_sequential": bool(same), "depth_tuning_s": round(tune_s, 2),
            "batched_host_split_s": {k: round(v, 4) for k, v in bat.timing.items()},
            "batched_cycle_all_active": {"instances_active": all_active, "ms": round(batch_cycle_s * 1e3, 3),
                                         "smoothing_steps_per_launch": m_b,
                                         "algorithmic_gb": round(batch_bytes / 1e9, 3),
                                         "gbs": round(batch_bytes / batch_cycle_s / 1e9, 1),
                                         "hbm_frac": round(batch_bytes / batch_cycle_s / 1e9 / HBM_PEAK_GBS, 4)},
            "single_instance_cycle": {"n": specs[big]["n"], "levels": levels[big], "ms": round(one_cycle_s * 1e3, 3),
                                      "gbs": round(one_bytes / one_cycle_s / 1e9, 1)},
            "batch_smooth_rows": {str(k): v for k, v in bat.batch.smooth_rows.items()},
            "per_instance": [[s_["n"], L, [int(c) for c in p.inner_counts]] for s_, L, p in zip(specs, levels, seq)]}


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
                         "gradient_s": round(t_bwd, 4), "gradient_over_forward": round(t_bwd / t_fwd, 3),
                         "adjoint_kernel_us": round(kern * 1e6, 2),
                         "adjoint_kernel_gbs": round(adj_bytes / kern / 1e9, 1),
                         "adjoint_bytes_per_cell": 26 * S + 32,
                         "grad_norm": float(mu.grad.double().norm()), "misfit": float(J)}
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
                    "forward_s": round(t_fwd, 4), "gradient_s": round(t_bwd, 4),
                    "gradient_over_forward": round(t_bwd / t_fwd, 3),
                    "adjoint_kernel_us": round(kern * 1e6, 2), "adjoint_kernel_gbs": round(adj_bytes / kern / 1e9, 1),
                    "adjoint_bytes_per_cell": 28 * S + 16,
                    "grad_norm": float(Kg.grad.double().norm()), "misfit": float(Jd)}
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
            "misfit": float(J), "grad_K_norm": float(Kg.grad.double().norm()),
            "grad_mu_norm": float(mug.grad.double().norm()), "finite": finite}


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


def cpu_baseline(n, precision, seconds):
    import torch
    import problems as P
    threads = os.cpu_count() or 1
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


def only(torch, args):
    """Counter-profile entry: one kernel family, ``--steps`` launches."""
    n = 4096
    if args.only == "copy":
        # plain device copy of a known size: 1 GiB read + 1 GiB written per launch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy: {args.steps} launches of {a.numel() * 8} B")
    elif args.only == "c2":
        from src import autotune
        r = bench_darcy(torch, args.n, "f64/f64", args.steps, 2, chunk=16, tile_steps=args.tile_steps,
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
        print(json.dumps(bench_c5(torch, n_inst=args.instances)), flush=True)
    elif args.only == "adjoint":
        print(json.dumps(bench_adjoint(torch)), flush=True)
    elif args.only == "biot_adjoint":
        print(json.dumps(bench_biot_adjoint(torch, n=args.n, steps=args.steps)), flush=True)
    elif args.only == "large_elast":
        r = bench_large_elasticity(torch, n, args.steps, 2)
        log(f"elasticity {n}^2: {r['steps_per_launch']} steps per launch, {r['kernel_s'] * 1e6:.1f} us/launch, "
            f"{r['kernel_s'] * 1e6 / r['steps_per_launch']:.1f} us/step, {r['bytes'] / r['kernel_s'] / 1e9:.0f} GB/s")
    else:
        r = bench_darcy(torch, n - 2, "f64/f64", args.steps, 2, chunk=16, tile_steps=args.tile_steps,
                        tile_threads=args.tile_threads)
        log(f"darcy {n}^2: {r['launch_kind']} launches of {r['steps_per_launch']} steps "
            f"({r['tile_threads']} threads per tile), "
            f"{r['kernel_s'] * 1e6:.1f} us/launch, {r['kernel_s'] * 1e6 / r['steps_per_launch']:.1f} us/step, "
            f"{r['bytes'] / r['kernel_s'] / 1e9:.0f} GB/s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--instances", type=int, default=64, help="instances of the batched Biot extra")
    ap.add_argument("--tile-steps", type=int, default=0,
                    help="with --only c2 or large_darcy: steps per launch of the LDS-tile kernel instead of the tuned launch")
    ap.add_argument("--tile-threads", type=int, default=0, choices=[0, 256, 512, 704],
                    help="with --tile-steps: threads per tile block, which fixes the tile shape (256, 512: "
                         "32 x 32; 704: 44 x 48; 0: the library's choice)")
    ap.add_argument("--only", choices=["large_elast", "large_darcy", "copy", "c2", "c5", "adjoint", "biot_vcycle",
                                       "biot_vcycle_eager", "biot_adjoint", "c1"],
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

    log(f"darcy C2 {args.n}x{args.n} fp64: warmup {args.warmup}, steps {args.steps}")
    r = bench_darcy(torch, args.n, "f64/f64", args.steps, args.warmup)
    value = r["cells"] * args.steps / r["elapsed"] / 1e6
    achieved = r["bytes"] / r["kernel_s"] / 1e9
    from src import autotune
    d = r["solver"]
    kind, m = r["launch_kind"], r["steps_per_launch"]
    if kind == "tile":
        rows = None
        blocks = r["tile_blocks"]
        tx_, ty_ = autotune.TILE_SHAPES[r["tile_threads"]]
        kname = f"k_tile_steps<double, xb::DarcyColl<double>, 2, {tx_}, {ty_}, {r['tile_threads']}>"
    else:
        two = kind == "sweep2"
        rows = autotune.lookup(d._tune_key(two))
        strip = 60 if two else 62
        blocks = (-(-(-(-d.ny // strip) * -(-d.nx // rows)) // 4)) if rows else None
        kname = ("k_sweep2<double, xb::DarcyColl<double>, 2>" if two
                 else "k_sweep<double, xb::DarcyColl<double>, 0, 2, false, false>")
    traffic = profiled_traffic(kname, blocks, r["cells"]) if blocks else None
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["elapsed"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded log-normal permeability (variance 1, correlation 16 cells, seed 0)",
        "config": {"workload": f"C2 plain Darcy D2Q9 {args.n}x{args.n} (+ghost layers: {args.n + 2}x{args.n + 2}), "
                               "FP64, p=1 left / p=0 right, no-flux top/bottom",
                   "nx": args.n + 2, "ny": args.n + 2, "precision": "f64/f64", "graph_chunk": 64},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel and launch grid, "
                                        "calibrated on a copy (profiles/darcy_tile_fp64_hbm_traffic.md)")
                     if traffic else "not profiled for this launch grid",
                     "kernel": kname, "launch_kind": kind, "launch_blocks": blocks, "rows_per_wavefront": rows,
                     "threads_per_block": r["tile_threads"],
                     "kernel_us": round(r["kernel_s"] * 1e6, 3), "steps_per_launch": r["steps_per_launch"],
                     "bytes_per_cell_per_launch": 152, "cells": r["cells"],
                     "per_step_equivalent_gbs": round(achieved * r["steps_per_launch"], 1)},
    }
    extras = {}

    def extra(name, fn):
        # an extra that fails is reported, and the headline line is still printed
        try:
            extras[name] = fn()
        except Exception as e:  # noqa: BLE001
            log(f"extra {name} failed: {type(e).__name__}: {e}")
            extras[name] = {"error": f"{type(e).__name__}: {e}"}

    if not args.no_extras:
        log("darcy C2 fp32")
        r32 = bench_darcy(torch, args.n, "f32/f32", args.steps // 2, args.warmup // 2)
        extras["c2_fp32_mlups"] = round(r32["cells"] * (args.steps // 2) / r32["elapsed"] / 1e6, 1)
        extras["c2_fp32_kernel_gbs"] = round(r32["bytes"] / r32["kernel_s"] / 1e9, 1)
        log("C1 64^2 to 1e-10")
        extra("c1_elasticity_64_fp64", lambda: bench_c1(torch))
        log("C3 Biot 256^2 MG (4 levels, lambda = mu = 1), inner 1e-8")
        extra("c3_biot_256_mg4_fp64", lambda: bench_c3(torch, steps=5, max_cycles=60))
        lam, mu = 0.11 * 0.8 / (1 - 0.8 ** 2), 0.5 * 0.11 / 1.8
        log("Biot 256^2 MG (8 levels, material of the reference periodic driver), inner 1e-8")
        extra("c3_biot_256_mg8_ref_material_fp64",
              lambda: bench_c3(torch, levels=8, steps=10, lam=lam, mu=mu, K=0.1))
        log(f"batched Biot: {args.instances} instances vs one after another")
        extra("c5_batched_biot_fp64", lambda: bench_c5(torch, n_inst=args.instances))
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
            rl = bench_large_elasticity(torch, n, 200, 20)
            gbs = rl["bytes"] / rl["kernel_s"] / 1e9
            extras[f"elasticity_{n}_fp64"] = {"mlups": round(rl["cells"] * 200 / rl["elapsed"] / 1e6, 1),
                                              "kernel_us": round(rl["kernel_s"] * 1e6, 2), "gbs": round(gbs, 1),
                                              "steps_per_launch": rl["steps_per_launch"],
                                              "threads_per_block": rl["tile_threads"],
                                              "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 144}
            log(f"darcy {n}^2 fp64 (HBM-resident)")
            rd = bench_darcy(torch, n - 2, "f64/f64", 200, 20, chunk=32)
            gbs = rd["bytes"] / rd["kernel_s"] / 1e9
            extras[f"darcy_{n}_fp64"] = {"mlups": round(rd["cells"] * 200 / rd["elapsed"] / 1e6, 1),
                                         "kernel_us": round(rd["kernel_s"] * 1e6, 2), "gbs": round(gbs, 1),
                                         "steps_per_launch": rd["steps_per_launch"],
                                         "launch_kind": rd["launch_kind"], "threads_per_block": rd["tile_threads"],
                                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 152}
            if rd["launch_kind"] == "tile":
                tx_, ty_ = autotune.TILE_SHAPES[rd["tile_threads"]]
                kn = f"k_tile_steps<double, xb::DarcyColl<double>, 2, {tx_}, {ty_}, {rd['tile_threads']}>"
                extras[f"darcy_{n}_fp64"]["traffic"] = profiled_traffic(kn, rd["tile_blocks"], rd["cells"])
    line["extra"] = extras
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_seconds:.0f} s)")
        line["cpu_baseline"] = cpu_baseline(args.n, "f64/f64", args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
