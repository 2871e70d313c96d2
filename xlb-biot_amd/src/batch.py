"""Batched solves: many independent problem instances advanced together.

The reference solves one problem per solver object, and its parameter studies run
the drivers one after another.  Here the multigrid V-cycles of many
:class:`~src.elasticity.LBMElasticity` instances (any grid sizes, Lame parameters,
level counts, coupling forces) run as ONE sequence of batched launches: every
sweep, prolongation and fill of a V-cycle is a single kernel launch over the jobs
of all instances (``xb_batch_launch``), and the per-instance convergence test
(``xb_batch_converge``) runs on the device at the end of each cycle and clears the
activity flag of every instance that has converged, so converged instances stop
costing anything but a flag load per block.  A cycle of all instances is captured
once into a HIP graph per ping-pong parity; the host only polls the count of
active instances every few cycles.

Per instance the operations, their order and their operands are those of
``LBMElasticity._vcycle_nl`` / ``_vcycle`` (reference src/elasticity.py:1221-1312)
and the stopping rule is that of the residual-driven inner loop of the Biot step
(``VCycleGraph.run_to`` without patience), so a batched solve gives the same
populations, bit for bit, as the instances solved one after another.

:class:`BiotBatch` drives many :class:`~src.poroelasticity.LBMPoroelasticity`
instances through their physical steps: the per-step glue of each instance
(sources, gradients, Darcy step: poroelasticity.py:303-339) is the instance's own
code, and the elasticity iterations of all instances run as one batched solve.
"""
import ctypes
import time

import numpy as np
import torch

from . import _native as nat
from . import autotune
from .base import graph_capture


class BatchPlan:
    """One prepared batched launch: the job table lives in a device workspace."""

    _JOB_TYPES = {nat.XB_BATCH_ELAST_STEP: nat.ElastArgs, nat.XB_BATCH_PROLONG_DIV: nat.ProlongJob,
                  nat.XB_BATCH_FILL: nat.FillJob, nat.XB_BATCH_NORM: nat.NormJob,
                  nat.XB_BATCH_ELAST_SMOOTHN: nat.ElastArgs}

    def __init__(self, kind, dtype, jobs, inst, device, tag=None):
        # tag: (role, level[, steps]) of the launch in a V-cycle schedule
        self.tag = tag
        lib = nat.lib()
        n = len(jobs)
        if n == 0:
            raise ValueError("a batch plan needs at least one job")
        arr = (self._JOB_TYPES[kind] * n)(*jobs)
        ids = (ctypes.c_int32 * n)(*inst)
        code = nat.dtype_code(dtype)
        nbytes = int(lib.xb_batch_table_bytes(kind, code, arr, n))
        if nbytes < 0:
            nat.check(nat.XB_EINVAL, "xb_batch_table_bytes")
        self.table = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.plan = nat.BatchPlan()
        nat.check(lib.xb_batch_prepare(kind, code, arr, ids, n, nat.ptr(self.table), nbytes, ctypes.byref(self.plan),
                                       nat.stream()), "xb_batch_prepare")
        self._ref = ctypes.pointer(self.plan)
        self.kind, self.njobs = kind, n
        self._self_dev = None

    def launch(self, active):
        nat.check(nat.lib().xb_batch_launch(self._ref, nat.ptr(active), nat.stream()), "xb_batch_launch")

    def use_list(self, on=True):
        """Launch over the work list written by the last compaction of this plan."""
        self.plan.use_list = int(bool(on))

    def compact(self, active):
        """Rebuild this plan's work list from ``active`` (one launch)."""
        if self._self_dev is None:
            self._self_dev = plans_on_device([self], self.table.device)
        nat.check(nat.lib().xb_batch_compact(nat.ptr(self._self_dev), 1, nat.ptr(active), nat.stream()),
                  "xb_batch_compact")

    def norms(self, out):
        nat.check(nat.lib().xb_batch_norms(self._ref, nat.ptr(out), nat.stream()), "xb_batch_norms")

    def converge(self, ninst, state, tol, max_cycles):
        nat.check(nat.lib().xb_batch_converge(self._ref, ninst, ctypes.byref(state), float(tol), int(max_cycles),
                                              nat.stream()), "xb_batch_converge")


def plans_on_device(plans, device):
    """A device copy of the plans' structs (for ``xb_batch_compact``); the copy is
    taken now, so set ``use_list`` first.  Not inside a graph capture."""
    from .base import no_capture
    no_capture("a device array of batch plans")
    arr = (nat.BatchPlan * len(plans))(*[p.plan for p in plans])
    host = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    return host.to(device)


class _Instance:
    """Buffers of one instance: level-0 ping-pong state and smoothing temporaries,
    coarse-level state pairs and restricted residuals, divergence fields and the
    finest-level residual partials."""

    def __init__(self, el, nlf, fused):
        self.el, self.nlf, self.fused = el, nlf, fused
        dev, dt = el.device, el._tdt
        self.levels = el.levels
        nx, ny = el.grid_size(0)
        z = lambda c, shape: torch.zeros((c,) + shape, dtype=dt, device=dev)
        self.X = [z(8, (nx, ny)) for _ in range(2)]
        self.T = [z(8, (nx, ny)) for _ in range(2)]
        self.dX = [z(1, (nx, ny)) for _ in range(2)]
        self.dT = [z(1, (nx, ny)) for _ in range(2)]
        self.ext = z(2, (nx, ny))
        self.C, self.D, self.R, self.P = {}, {}, {}, {}
        for level in range(1, self.levels):
            cnx, cny = el.grid_size(level)
            x0, x1, y0, y1 = el.interior_box(level - 1)
            self.C[level] = [z(8, (cnx, cny)) for _ in range(2)]
            self.D[level] = [z(1, (cnx, cny)) for _ in range(2)]
            self.R[level] = z(8, ((x1 - x0) // 2, (y1 - y0) // 2))
        self.partial_count = {}

    def partials(self, level, count):
        ws = self.P.get(level)
        if ws is None or ws.numel() < count:
            ws = self.P[level] = torch.zeros(max(count, 1), dtype=torch.float64, device=self.el.device)
        self.partial_count[level] = count
        return ws


class ElasticityBatch:
    """Multigrid V-cycles of many elasticity solvers run as batched launches.

    ``solvers``: :class:`LBMElasticity` instances with ``MG_levels >= 2``, one
    precision and one ``MG_n_smoothing``; ``nonlinear_forces``: per instance
    ``None`` or a fused coupling force (``BiotNonlinearForce``, all or none).
    :meth:`solve` iterates every instance until its finest-level residual norm is
    at most ``tol`` times its scale, or ``max_cycles`` cycles.  ``scale``: "cycle"
    (the first cycle's residual) or "rhs" (the residual of the zero state under the
    current forces, ``|S(0)|``, computed on the device before the cycles)."""

    def __init__(self, solvers, nonlinear_forces=None, tol=1e-8, max_cycles=10000, poll=4, external_force=True,
                 scale="cycle", smooth_rows="tune", work_lists=True):
        if scale not in ("cycle", "rhs"):
            raise ValueError("scale must be 'cycle' or 'rhs'")
        self.scale = scale
        if not solvers:
            raise ValueError("an elasticity batch needs at least one solver")
        nlfs = list(nonlinear_forces) if nonlinear_forces is not None else [None] * len(solvers)
        if len(nlfs) != len(solvers):
            raise ValueError("one nonlinear force (or None) per solver")
        dt = solvers[0]._tdt
        ks = solvers[0].k_smooth
        for el in solvers:
            if el.levels < 2:
                raise ValueError("batched V-cycles need MG_levels >= 2 for every instance")
            el.precisionPolicy.require_single("the batched solve")
            if el._tdt != dt:
                raise ValueError("all instances of a batch share one precision")
            if el.k_smooth != ks:
                raise ValueError("all instances of a batch share MG_n_smoothing")
            if el._lam_field is not None:
                raise NotImplementedError("batched V-cycles take scalar Lame parameters per instance")
        fused = [getattr(n, "fused", False) for n in nlfs]
        if any(n is not None for n in nlfs) and not all(fused):
            raise ValueError("batched coupling forces must all be fused BiotNonlinearForce objects")
        self.fused = all(fused) and nlfs[0] is not None
        # finest-level external force buffers (zero unless given to solve())
        self.external_force = bool(external_force)
        self.tol, self.max_cycles, self.poll = float(tol), int(max_cycles), max(1, int(poll))
        self.dtype, self.k_smooth = dt, ks
        self.device = solvers[0].device
        self.inst = [_Instance(el, nlf, self.fused) for el, nlf in zip(solvers, nlfs)]
        n = len(self.inst)
        self.n = n
        dev = self.device
        self.r0 = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
        self.rlast = torch.zeros(n, dtype=torch.float64, device=dev)
        self.cycles = torch.zeros(n, dtype=torch.int32, device=dev)
        self.active = torch.ones(n, dtype=torch.int32, device=dev)
        self.status = torch.zeros(n, dtype=torch.int32, device=dev)
        self.nactive = torch.zeros(1, dtype=torch.int32, device=dev)
        # reference residual per instance (<= 0: none), see solve(r_ref=...)
        self.rref = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
        self.state = nat.ConvergeState(nat.ptr(self.r0), nat.ptr(self.rlast), nat.ptr(self.cycles),
                                       nat.ptr(self.active), nat.ptr(self.status), nat.ptr(self.nactive),
                                       nat.ptr(self.rref))
        # launch shape of the batched smoothing sweeps per level: the single-grid
        # choice favours many short wavefront walks (parallelism on one small grid),
        # a batch has parallelism to spare and prefers longer walks (fewer halo
        # re-reads, fewer blocks).  Residual sweeps keep the single-grid shape so that
        # their norm partials, hence the stopping decisions, are those of the
        # instances solved one by one.
        L = max(I.levels for I in self.inst)
        if smooth_rows == "tune":
            self.smooth_rows = self._tune_smooth_rows(L)
        elif isinstance(smooth_rows, dict):
            self.smooth_rows = dict(smooth_rows)
        else:
            self.smooth_rows = {lvl: smooth_rows for lvl in range(L)} if smooth_rows else {}
        self.schedules = [self._schedule(src) for src in (0, 1)]
        # Work lists: a cycle of the listed form first lists, per launch, the jobs of
        # the instances still active, and every launch then runs a grid of as many
        # workgroups as the GPU holds at once over their blocks only (no workgroup is
        # dispatched for a converged instance).  With most of the work active, one
        # workgroup per block is faster (measured on the level-0 smoothing tiles of the
        # 64-instance set: 212 us per launch against 236 us with every instance active,
        # 82 us against 45 us with an eighth of them); a solve replays the cycle of one
        # workgroup per block until the active instances hold less than
        # ``list_switch`` of the finest-level cells, then the listed cycle (0.9: on the
        # 64-instance set, 4 physical steps took 0.806-0.816 s against 0.835-0.844 s
        # with 0.7, 0.819-0.828 s with the listed cycle throughout, at 2-8 cycles per poll).
        self.work_lists = bool(work_lists)
        self.list_switch = 0.9
        self._plans_dev = None
        if self.work_lists:
            for sched in self.schedules:
                for p in sched:
                    p.use_list(True)
            self._plans_dev = [plans_on_device(sched, self.device) for sched in self.schedules]
        self._cells0 = np.array([float(I.el.nx * I.el.ny) for I in self.inst])
        self._rhs_plan = self._rhs_schedule() if scale == "rhs" else None
        # eager cycles (buffers, tables and tuned shapes settle), then capture
        self._reset()
        self._run_schedule(0, listed=False)
        if self.work_lists:
            self._reset()
            self._run_schedule(0, listed=True)
        torch.cuda.synchronize()
        self.graphs = self._capture_cycles(listed=False)
        self.list_graphs = self._capture_cycles(listed=True) if self.work_lists else None
        torch.cuda.synchronize()

    def _capture_cycles(self, listed):
        graphs = []
        for src in (0, 1):
            g = torch.cuda.CUDAGraph()
            with graph_capture(g):
                self._run_schedule(src, listed)
            graphs.append(g)
        return graphs

    # ---- schedule of one V-cycle of every instance (ping-pong parity src)
    def _sweep_args(self, I, level, mode, f_in, f_out, div_in, div_out, coarse=None, partials=None, rows=None):
        el = I.el
        ext = I.ext if level == 0 and self.external_force else None
        resid = I.R[level] if level > 0 else None
        a = el._launch(f_in, f_out, level, mode, ext=ext, resid=resid, coarse=coarse)
        if partials is not None:
            a.norm_partials = nat.ptr(partials)
        if self.fused:
            el._nl_args(a, div_in, div_out, I.nlf, level)
        if rows is None and mode == nat.XB_MODE_SMOOTH:
            rows = getattr(self, "smooth_rows", {}).get(level)
        if rows:
            a.rows = int(rows)
        return a

    def _tile_steps(self, level):
        """Smoothing steps per batched launch at ``level``: the single-instance choice
        (``LBMElasticity._smooth_tile_steps``) when every instance at that level agrees,
        else one sweep per step."""
        ms = {I.el._smooth_tile_steps(level, I.nlf if self.fused else None)
              for I in self.inst if I.levels > level}
        if not self.fused:
            return 1
        return ms.pop() if len(ms) == 1 else 1

    def _signature(self):
        import hashlib
        sig = repr(sorted((I.el.nx, I.el.ny, I.levels, I.el._tables.get(0) is not None) for I in self.inst))
        return hashlib.sha1(f"{sig}/{self.dtype}/{self.fused}/{self.external_force}".encode()).hexdigest()[:16]

    def _tune_smooth_rows(self, L, candidates=autotune.CANDIDATES, reps=5):
        """Launch shape of each level's batched smoothing sweep: every candidate is
        timed on the whole batch with HIP events; the fastest is kept and cached."""
        out = {}
        sig = self._signature()
        active = torch.ones(self.n, dtype=torch.int32, device=self.device)
        for level in range(L):
            key = f"batch/elast_smooth/{sig}/level{level}"
            rows = autotune.lookup(key)
            if rows:
                out[level] = rows
                continue
            ids = [i for i, I in enumerate(self.inst) if I.levels > level]

            def run(r, level=level, ids=ids):
                jobs = []
                for i in ids:
                    I = self.inst[i]
                    f, o, d, od = ((I.X[0], I.T[0], I.dX[0], I.dT[0]) if level == 0 else
                                   (I.C[level][0], I.C[level][1], I.D[level][0], I.D[level][1]))
                    jobs.append(self._sweep_args(I, level, nat.XB_MODE_SMOOTH, f, o, d, od, rows=r))
                plan = BatchPlan(nat.XB_BATCH_ELAST_STEP, self.dtype, jobs, ids, self.device)
                return autotune.time_launch(lambda: plan.launch(active), reps)
            timings = {r: run(r) for r in candidates}
            best = min(timings, key=timings.get)
            autotune.store(key, best, timings)
            out[level] = best
        return out

    def _schedule(self, src):
        """A list of ``BatchPlan`` objects (and the final norm plan) for one cycle."""
        plans = []
        dt = self.dtype
        L = max(I.levels for I in self.inst)
        ks = self.k_smooth
        cur = {}
        for i, I in enumerate(self.inst):
            cur[i, 0] = (I.X[src], I.dX[src])

        def next_buf(i, level, last_post):
            I = self.inst[i]
            f, _ = cur[i, level]
            if level == 0:
                if last_post:
                    return I.X[1 - src], I.dX[1 - src]
                k = 1 if f is I.T[0] else 0
                return I.T[k], I.dT[k]
            k = 1 if f is I.C[level][0] else 0
            return I.C[level][k], I.D[level][k]

        def smooth(level, last_step, m=1, coarse=None):
            # m steps in one launch: on LDS tiles (m > 1) or one sweep; coarse (tile
            # launches only): instance -> the next level's correction, injected first
            jobs, ids = [], []
            for i, I in enumerate(self.inst):
                if I.levels <= level:
                    continue
                f, d = cur[i, level]
                of, od = next_buf(i, level, last_step and level == 0)
                # f None: a coarse level's zero start state (tile launches only; the
                # divergence input is then not read and od marks the coupling)
                a = self._sweep_args(I, level, nat.XB_MODE_SMOOTH, f, of, d if f is not None else od, od)
                if m > 1:
                    a.rows = m
                if coarse is not None and i in coarse:
                    a.coarse_out = nat.ptr(coarse[i])
                jobs.append(a)
                ids.append(i)
                cur[i, level] = (of, od)
            if jobs:
                kind = nat.XB_BATCH_ELAST_SMOOTHN if m > 1 else nat.XB_BATCH_ELAST_STEP
                plans.append(BatchPlan(kind, dt, jobs, ids, self.device, tag=("smooth", level, m)))

        def smooth_phase(level, post, coarse=None):
            m = self._tile_steps(level)
            done = 0
            while done < ks:
                step = min(m, ks - done)
                done += step
                smooth(level, post and done == ks, step, coarse if done == step else None)

        norm_jobs = []
        for level in range(L):
            smooth_phase(level, False)
            deeper = [i for i, I in enumerate(self.inst) if I.levels - 1 > level]
            if not deeper:
                continue
            jobs, fills, fill_ids = [], [], []
            zero_start = ks > 1 and self._tile_steps(level + 1) > 1
            for i in deeper:
                I = self.inst[i]
                f, d = cur[i, level]
                a = self._sweep_args(I, level, nat.XB_MODE_RESIDUAL, f, None, d, None, coarse=I.R[level + 1])
                count = int(nat.lib().xb_elast_partials_count(a))
                ws = I.partials(level, count)
                a.norm_partials = nat.ptr(ws)
                jobs.append(a)
                if level == 0:
                    norm_jobs.append((nat.NormJob(nat.ptr(ws), count, 0), i))
                if zero_start:
                    # the first smoothing launch of the coarse level starts from zero
                    cur[i, level + 1] = (None, None)
                    continue
                c0, d0 = I.C[level + 1][0], I.D[level + 1][0]
                fills.append(nat.FillJob(c0.numel(), nat.ptr(c0), 0.0))
                fills.append(nat.FillJob(d0.numel(), nat.ptr(d0), 0.0))
                fill_ids += [i, i]
                cur[i, level + 1] = (c0, d0)
            plans.append(BatchPlan(nat.XB_BATCH_ELAST_STEP, dt, jobs, deeper, self.device, tag=("residual", level)))
            if fills:
                plans.append(BatchPlan(nat.XB_BATCH_FILL, dt, fills, fill_ids, self.device, tag=("fill", level + 1)))
        for level in range(L - 1, -1, -1):
            deeper = [i for i, I in enumerate(self.inst) if I.levels - 1 > level]
            if deeper and ks > 1 and self._tile_steps(level) > 1:
                # the prolongation rides in the first post-smoothing tile launch
                smooth_phase(level, True, {i: cur[i, level + 1][0] for i in deeper})
                continue
            if deeper:
                jobs = []
                for i in deeper:
                    I, el = self.inst[i], self.inst[i].el
                    f, d = cur[i, level]
                    fc, _ = cur[i, level + 1]
                    nx, ny = el.grid_size(level)
                    cnx, cny = el.grid_size(level + 1)
                    x0, x1, y0, y1 = el.interior_box(level)
                    jobs.append(nat.ProlongJob(nx, ny, cnx, cny, x0, x1, y0, y1, el.lamda, el.mu, nat.ptr(f),
                                               nat.ptr(fc), nat.ptr(d) if self.fused else None))
                plans.append(BatchPlan(nat.XB_BATCH_PROLONG_DIV, dt, jobs, deeper, self.device, tag=("prolong", level)))
            smooth_phase(level, True)
        if len(norm_jobs) != self.n:
            raise RuntimeError("every instance needs a finest-level residual")
        norm_jobs.sort(key=lambda t: t[1])
        self._norm_plan = BatchPlan(nat.XB_BATCH_NORM, dt, [j for j, _ in norm_jobs], [i for _, i in norm_jobs],
                                    self.device)
        return plans

    def _rhs_schedule(self):
        """Finest-level residual sweep of the zero state of every instance (its
        partials feed the norm plan, its restriction goes to scratch)."""
        jobs = []
        for I in self.inst:
            nx, ny = I.el.grid_size(0)
            I.Z = torch.zeros((8, nx, ny), dtype=self.dtype, device=self.device)
            I.dZ = torch.zeros((1, nx, ny), dtype=self.dtype, device=self.device)
            a = self._sweep_args(I, 0, nat.XB_MODE_RESIDUAL, I.Z, None, I.dZ, None, coarse=I.R[1])
            a.norm_partials = nat.ptr(I.P[0])
            if int(nat.lib().xb_elast_partials_count(a)) != I.partial_count[0]:
                raise RuntimeError("zero-state residual and cycle residual disagree in shape")
            jobs.append(a)
        return BatchPlan(nat.XB_BATCH_ELAST_STEP, self.dtype, jobs, list(range(self.n)), self.device)

    def _run_schedule(self, src, listed):
        """One V-cycle of every active instance: one workgroup per table block, or
        (``listed``) over the work lists the cycle builds first."""
        if listed:
            nat.check(nat.lib().xb_batch_compact(nat.ptr(self._plans_dev[src]), len(self.schedules[src]),
                                                 nat.ptr(self.active), nat.stream()), "xb_batch_compact")
        for p in self.schedules[src]:
            p.use_list(listed)
            p.launch(self.active)
        self._norm_plan.converge(self.n, self.state, self.tol, self.max_cycles)

    def _listed_now(self, act):
        """Whether the active instances (host mask) hold less than ``list_switch`` of
        the finest-level cells."""
        return self.list_graphs is not None and \
            float(self._cells0[act != 0].sum()) < self.list_switch * float(self._cells0.sum())

    def _reset(self, active=None):
        self.r0.fill_(-1.0)
        self.rlast.zero_()
        self.cycles.zero_()
        self.status.zero_()
        if active is None:
            self.active.fill_(1)
        else:
            self.active.copy_(torch.as_tensor(active, dtype=torch.int32))
        self.nactive.fill_(int(self.active.sum().item()))

    # ---- solves
    def load(self, states, forces=None):
        """Copy initial populations (SoA ``(8, nx, ny)``) and external forces
        (``(2, nx, ny)`` or None = zero) into the instances' static buffers."""
        for i, I in enumerate(self.inst):
            I.X[0].copy_(states[i])
            if forces is not None and forces[i] is not None:
                if not self.external_force:
                    raise ValueError("this batch was built without external forces")
                I.ext.copy_(forces[i])
            else:
                I.ext.zero_()
            if self.fused:
                el = I.el
                nat.check(nat.lib().xb_elast_divergence(el.nx, el.ny, nat.dtype_code(self.dtype), el.lamda, el.mu,
                                                        nat.ptr(I.X[0]), None, nat.ptr(I.dX[0]), None, None, nat.stream()),
                          "xb_elast_divergence")

    def solve(self, states, forces=None, active=None):
        """V-cycles of every (active) instance until each converges (residual at most
        ``tol`` times its scale).  Returns the list of final SoA states (views of the
        batch's buffers, valid until the next solve), and sets :attr:`last_cycles`,
        :attr:`last_rel`, :attr:`last_scale` and :attr:`last_status` (1 converged,
        2 out of cycles)."""
        self.load(states, forces)
        self._reset(active)
        self.rref.fill_(-1.0)
        if self._rhs_plan is not None:
            self._rhs_plan.launch(self.active)
            self._norm_plan.norms(self.rref)
        cur, done = 0, 0
        limit = self.max_cycles + 2 * self.poll
        act = self.active.cpu().numpy()
        graphs = self.list_graphs if self._listed_now(act) else self.graphs
        while done < limit:
            for _ in range(self.poll):
                graphs[cur].replay()
                cur = 1 - cur
                done += 1
            # the activity flags after these cycles (one read: it synchronises)
            act = self.active.cpu().numpy()
            if not act.any():
                break
            if graphs is self.graphs and self._listed_now(act):
                graphs = self.list_graphs
        cycles = self.cycles.cpu().numpy()
        status = self.status.cpu().numpy()
        r0 = self.r0.cpu().numpy()
        rl = self.rlast.cpu().numpy()
        bad = [i for i in range(self.n) if status[i] == 3]
        if bad:
            raise FloatingPointError(f"V-cycle residual is not finite for instances {bad}")
        self.last_cycles = cycles
        self.last_status = status
        ref = self.rref.cpu().numpy()
        scale = np.where(ref > 0, ref, r0)
        self.last_scale = scale
        with np.errstate(invalid="ignore", divide="ignore"):
            self.last_rel = np.where(scale > 0, rl / np.where(scale > 0, scale, 1.0), 0.0)
        return [I.X[int(cycles[i]) % 2] for i, I in enumerate(self.inst)]


class BiotBatch:
    """Physical steps of many Biot problems with one batched elasticity solve per
    step.  ``problems``: :class:`LBMPoroelasticity` instances with multigrid
    elasticity (``elasticity_MG_levels >= 2``) and the fused coupling force; the
    inner tolerance is each problem's ``elasticity_tol``, which must agree."""

    def __init__(self, problems, poll=4):
        if not problems:
            raise ValueError("a Biot batch needs at least one problem")
        tols = {p.elasticityTol for p in problems}
        if len(tols) != 1 or None in tols:
            raise ValueError("every problem needs the same elasticity_tol")
        maxs = {p.elasticityMaxSteps for p in problems}
        if len(maxs) != 1:
            raise ValueError("every problem needs the same elasticity_max_steps")
        for p in problems:
            if not p.fusedCoupling:
                raise ValueError("batched Biot steps use the fused coupling force")
        scales = {p.elasticityTolScale for p in problems}
        if len(scales) != 1:
            raise ValueError("every problem needs the same elasticity_tol_scale")
        self.problems = problems
        self.batch = ElasticityBatch([p.elasticity for p in problems], [p.coupling_force() for p in problems],
                                     tol=tols.pop(), max_cycles=maxs.pop(), poll=poll, scale=scales.pop())
        self.cycles = []
        # host seconds spent in the per-instance step glue and in the batched solves
        self.timing = {"pre": 0.0, "solve": 0.0, "post": 0.0}

    def initial_state(self):
        return [p.initial_state() for p in self.problems]

    def step(self, states, timestep):
        """One physical step of every problem: ``states`` is a list of
        ``(f_el, f_darcy, dt_div_eta)`` SoA triples; returns the list of 7-tuples of
        ``LBMPoroelasticity._step_soa``."""
        t0 = time.perf_counter()
        ctxs = [p._pre_inner(fe, fd, dde, timestep) for p, (fe, fd, dde) in zip(self.problems, states)]
        t1 = time.perf_counter()
        fes = self.batch.solve([s[0] for s in states], [c["force"] for c in ctxs])
        t2 = time.perf_counter()
        self.cycles.append(self.batch.last_cycles.copy())
        out = []
        for i, (p, c) in enumerate(zip(self.problems, ctxs)):
            p.inner_counts.append(int(self.batch.last_cycles[i]))
            p.inner_residuals.append(float(self.batch.last_rel[i]))
            out.append(p._post_inner(c, fes[i].clone()))
        t3 = time.perf_counter()
        self.timing["pre"] += t1 - t0
        self.timing["solve"] += t2 - t1
        self.timing["post"] += t3 - t2
        return out

    def run(self, t_max):
        """``t_max + 1`` physical steps of every problem (as ``LBMPoroelasticity.run``
        without I/O); returns the list of ``(f_el, f_darcy)`` SoA pairs."""
        states = self.initial_state()
        for t in range(t_max + 1):
            outs = self.step(states, t)
            states = [(o[0], o[1], o[4]) for o in outs]
        return [(s[0], s[1]) for s in states]
