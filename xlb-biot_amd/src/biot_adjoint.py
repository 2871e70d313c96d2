"""Reverse-mode gradients through the coupled Biot step with multigrid elasticity.

:func:`biot_solve` runs ``steps`` coupled steps of an :class:`LBMPoroelasticity`
(reference src/poroelasticity.py:303-339), each with the problem's
``elasticity_steps`` V-cycles (src/elasticity.py:1221-1312) and the fused coupling
force, from the problem's initial state.  It returns the displacement ``eta``
``(2, nx, ny)`` and the pressure ``(1, nx, ny)`` of the last step (SoA), and is
differentiable with respect to the Darcy relaxation-rate field omega(x) (the
permeability through omega = 1 / (3 K + 1/2)) and the Lame fields lambda(x), mu(x)
of the finest elasticity grid.

Backward is the discrete adjoint of exactly that computation, what ``jax.grad``
through the reference's steps would give:

* the forward pass checkpoints the inputs of a physical step every
  ``checkpoint_every`` V-cycles (pseudo-steps) - every ``checkpoint_every //
  elasticity_steps`` steps, and inside steps of more V-cycles than that, their
  V-cycle inputs every ``checkpoint_every`` V-cycles;
* backward recomputes one segment at a time from its checkpoint (a recomputed
  V-cycle keeps the inputs of its finest-level sweeps), then reverses it, physical
  step by physical step and V-cycle by V-cycle;
* a V-cycle is reversed level by level: transposed smoothing and residual sweeps
  (``xb_elast_mg_adjoint``; the residual's 2x2 restriction is transposed inside),
  the transposed coupling force of every sweep (``xb_elast_coupling_adjoint``) and
  the transposed prolongation (``xb_elast_restrict_sum``);
* the coupler's pointwise maps and lattice gradients are transposed by
  ``xb_elast_macro_adjoint``, ``xb_elast_divergence_adjoint``,
  ``xb_darcy_macro_adjoint`` and ``xb_darcy_gradient_adjoint``, the Darcy step by
  ``xb_darcy_adjoint``.

Parameter derivatives and the adjoint of each step's static force accumulate in
FP64.  Coarse multigrid levels use the scalar lambda, mu (they carry no parameter
derivative); the initial state is taken as given (not differentiated).
"""
import ctypes

import torch

from . import _native as nat
from .base import graph_capture


def _zeros(shape, dtype, device):
    return torch.zeros(shape, dtype=dtype, device=device)


class BiotAdjointSession:
    """Buffers, checkpoints and kernels of the differentiable coupled solve of one
    problem (kept on the problem and reused by later calls of the same shape)."""

    def __init__(self, problem, steps, K):
        p = problem
        el, da = p.elasticity, p.darcy
        if el.levels < 2:
            raise ValueError("biot_solve differentiates multigrid elasticity (elasticity_MG_levels >= 2)")
        if p.elasticityTol is not None:
            raise ValueError("biot_solve runs the fixed elasticity_steps V-cycles per step (elasticity_tol=None)")
        if not p.fusedCoupling:
            raise ValueError("biot_solve needs the fused coupling force (fused_coupling=True)")
        if steps < 1 or K < 1:
            raise ValueError("need steps >= 1 and a checkpoint interval >= 1")
        self.p, self.el, self.da = p, el, da
        self.steps, self.K, self.n = int(steps), int(K), p.elasticitySteps
        # physical steps per checkpoint: K V-cycles' worth (at least one step; a step
        # of more than K V-cycles also checkpoints its V-cycles every K)
        self.P = max(1, self.K // self.n)
        self.dtype, self.device = el._tdt, el.device
        nx, ny = el.nx, el.ny
        if el._lam_field is None:
            el.set_lame_fields(torch.full((nx, ny), el.lamda, dtype=torch.float64),
                               torch.full((nx, ny), el.mu, dtype=torch.float64))
        if da._omega_field is None:
            da.set_omega_field(torch.full((nx, ny), da.omega, dtype=torch.float64))
        self.nlf = p.coupling_force()
        self.vg = el.vcycle_graph(_zeros((2, nx, ny), self.dtype, self.device), self.nlf, keep_coarse=True)
        # the forward keeps every V-cycle's coarse correction (a quarter of a state) when
        # they fit this many bytes: the recompute then redoes only the finest level
        nxc, nyc = el.grid_size(1)
        fc_bytes = self.steps * self.n * 8 * nxc * nyc * torch.tensor([], dtype=self.dtype).element_size()
        self.keep_coarse = fc_bytes <= float(getattr(p, "adjoint_coarse_budget_bytes", 8e9))
        self.fc = []
        self.generation = 0
        self.ckpt_step = []
        self.ckpt_cycle = []
        # coarse grids of at most this many cells run their transposed smoothing in one
        # workgroup (xb_elast_smooth_adjoint_small, one launch per run of sweeps; None:
        # the grids it holds in LDS); default 0, the LDS tiles on every level.  C4
        # backward (1024 x 1026): tiles 0.350 s, single workgroup from LDS on the four
        # smallest levels 0.361 s, from global memory below 600 cells 0.378 s - the
        # per-cell wall-rule reads of the single workgroup stay in global memory
        small = getattr(p, "adjoint_small_cells", 0)
        self.adjoint_small_cells = None if small is None else int(small)
        # transposed smoothing sweeps per launch on LDS tiles (0: one launch pair per
        # sweep, xb_elast_mg_adjoint + xb_elast_coupling_adjoint): the problem's
        # adjoint_tile_steps on every level when set, else 2 on levels 0 and 1 and 3 below
        # them (the C4 backward: 0.348 against 0.355 s with 2 everywhere; the same gradients)
        self.adjoint_tile_steps = getattr(p, "adjoint_tile_steps", None)
        self._coarse_adjoint_graph()

    def _coarse_adjoint_graph(self):
        """The transposed coarse V-cycle (levels >= 1) as one captured HIP graph.  It
        is the same linear map in every V-cycle (no recorded states, no parameter
        derivatives): the restricted adjoint in :attr:`ab_c` to the adjoint of the
        level-0 residual's restriction in :attr:`rbar_c`."""
        el = self.el
        nxc, nyc = el.grid_size(1)
        x0, x1, y0, y1 = el.interior_box(0)
        self.ab_c = _zeros((8, nxc, nyc), self.dtype, self.device)
        self.rbar_c = _zeros((8, (x1 - x0) // 2, (y1 - y0) // 2), self.dtype, self.device)
        no_acc = {"lam": None, "mu": None, "base": None}

        def body():
            nat.fill(self.rbar_c, 0.0)
            self._level_adjoint(1, self.ab_c, None, self.rbar_c, no_acc)
        body()  # eager once: the gradient wall tables of the coarse grids exist before capture
        torch.cuda.synchronize()
        self.coarse_graph = torch.cuda.CUDAGraph()
        with graph_capture(self.coarse_graph):
            body()
        # the coarse solve of a recomputed V-cycle (residual of the pre-smoothed state,
        # its restriction, the coarse V-cycle from zero) from static inputs
        nx, ny = el.grid_size(0)
        self.cs_f = _zeros((8, nx, ny), self.dtype, self.device)
        self.cs_div = _zeros((1, nx, ny), self.dtype, self.device)
        self.cs_base = _zeros((2, nx, ny), self.dtype, self.device)
        el._coarse_solve_nl(self.cs_f, self.cs_div, 0, None, self.cs_base, self.nlf)
        torch.cuda.synchronize()
        self.solve_graph = torch.cuda.CUDAGraph()
        with graph_capture(self.solve_graph):
            self.cs_out, _ = el._coarse_solve_nl(self.cs_f, self.cs_div, 0, None, self.cs_base, self.nlf)

    # ---- parameters
    def bind(self, omega, lam, mu):
        el, da = self.el, self.da
        el.set_lame_fields(lam.detach().reshape(el.nx, el.ny), mu.detach().reshape(el.nx, el.ny))
        da.set_omega_field(omega.detach().reshape(da.nx, da.ny))

    # ---- forward
    def forward(self):
        self.generation += 1
        p, vg = self.p, self.vg
        fe, fd, dde = p.initial_state()
        self.ckpt_step, self.ckpt_cycle, self.fc = [], [], []
        out = None
        for t in range(self.steps):
            if t % self.P == 0:
                self.ckpt_step.append((fe.clone(), fd.clone(), dde.clone()))
            ctx = p._pre_inner(fe, fd, dde, t)
            vg.set_force(ctx["force"])
            vg.load(fe)
            cps, fcs = [], []
            for j in range(self.n):
                if j > 0 and j % self.K == 0:
                    cps.append(vg.state.clone())
                vg.replay()
                if self.keep_coarse:
                    fcs.append(vg.coarse.clone())
            self.ckpt_cycle.append(cps)
            self.fc.append(fcs)
            out = p._post_inner(ctx, vg.state.clone())
            fe, fd, dde = out[0], out[1], out[4]
        return out[2], out[5]

    # ---- recompute of V-cycles with the finest-level sweep inputs kept
    def _vcycle_record(self, f, div, base, f_c=None):
        """One V-cycle from ``f`` (divergence ``div``) with its finest-level sweep
        inputs kept; ``f_c``: its coarse correction when the forward kept it (else the
        coarse levels are solved again)."""
        el, nlf, k = self.el, self.nlf, self.el.k_smooth
        states = []
        for _ in range(k):
            states.append(f)
            f, div = el._smooth_nl(f, div, 0, None, base, nlf)
        states.append(f)
        if f_c is None:
            self.cs_f.copy_(f)
            self.cs_div.copy_(div)
            self.cs_base.copy_(base)
            self.solve_graph.replay()
            f_c = self.cs_out
        nxc, nyc = el.grid_size(1)
        f, div = f.clone(), div.clone()
        nx, ny = el.grid_size(0)
        x0, x1, y0, y1 = el.interior_box(0)
        nat.check(nat.lib().xb_elast_prolong_add_div(nx, ny, nat.dtype_code(f.dtype), el.lamda, el.mu, nat.ptr(f),
                                                     nat.ptr(f_c), nxc, nyc, x0, x1, y0, y1, nat.ptr(div),
                                                     *el._fields(), nat.stream()), "xb_elast_prolong_add_div")
        for _ in range(k):
            states.append(f)
            f, div = el._smooth_nl(f, div, 0, None, base, nlf)
        return f, div, states

    def _recompute(self, start, m, base, fcs=None):
        f = start
        div = self.el._divergence_soa(f)
        recs = []
        for i in range(m):
            f, div, states = self._vcycle_record(f, div, base, fcs[i] if fcs else None)
            recs.append(states)
        return recs, f

    # ---- transposed V-cycle
    def _sweep_adjoint(self, level, mode, ab_in, ab_add, state, rbar, acc):
        el = self.el
        nx, ny = el.grid_size(level)
        x0, x1, y0, y1 = el.interior_box(level)
        out = torch.empty((8, nx, ny), dtype=self.dtype, device=self.device)
        fbar = torch.empty((2, nx, ny), dtype=self.dtype, device=self.device)
        grads = level == 0 and state is not None and acc.get("lam") is not None
        fields = el._fields() if level == 0 else (None, None)
        table = el._tables.get(level)
        a = nat.ElastMGAdjointArgs(nx, ny, nat.dtype_code(self.dtype), mode, el.lamda, el.mu, el.gamma, fields[0],
                                   fields[1], x0, x1, y0, y1, nat.ptr(ab_in), nat.ptr(ab_add), nat.ptr(out),
                                   nat.ptr(rbar), nat.ptr(fbar), nat.ptr(state) if grads else None, None,
                                   nat.ptr(acc["lam"]) if grads else None, nat.ptr(acc["mu"]) if grads else None,
                                   table.ref() if table is not None else None)
        nat.check(nat.lib().xb_elast_mg_adjoint(a, nat.stream()), "xb_elast_mg_adjoint")
        gt = self.da._grad_table(nx, ny)
        nat.check(nat.lib().xb_elast_coupling_adjoint(
            nx, ny, nat.dtype_code(self.dtype), el.lamda, el.mu, fields[0], fields[1], nat.ptr(fbar),
            float(self.nlf.scale), gt.ref() if gt is not None else None, nat.ptr(out),
            nat.ptr(state) if grads else None, nat.ptr(acc["lam"]) if grads else None,
            nat.ptr(acc["mu"]) if grads else None, nat.ptr(acc["base"]) if level == 0 else None, nat.stream()),
            "xb_elast_coupling_adjoint")
        return out

    def _smooth_adjoint_run(self, level, ab, states, rbar, acc):
        """The level's ``k`` transposed smoothing sweeps, last sweep first (``states``:
        their recorded inputs, or ``None`` entries): on LDS tiles, up to
        ``adjoint_tile_steps`` sweeps per launch (``xb_elast_smooth_adjoint_tiles``),
        or a launch pair per sweep."""
        el = self.el
        S = nat.XB_ADJ_SMOOTH
        m = int(self.adjoint_tile_steps) if self.adjoint_tile_steps is not None else (3 if level >= 2 else 2)
        j = el.k_smooth - 1
        if m < 1:
            while j >= 0:
                ab = self._sweep_adjoint(level, S, ab, None, states[j], rbar, acc)
                j -= 1
            return ab
        nx, ny = el.grid_size(level)
        x0, x1, y0, y1 = el.interior_box(level)
        fields = el._fields() if level == 0 else (None, None)
        table = el._tables.get(level)
        gt = self.da._grad_table(nx, ny)
        base = acc.get("base") if level == 0 else None
        while j >= 0:
            step = min(m, j + 1)
            recs = [states[j - s] for s in range(step)]
            grads = level == 0 and recs[0] is not None and acc.get("lam") is not None
            out = torch.empty((8, nx, ny), dtype=self.dtype, device=self.device)
            a = nat.ElastMGAdjointArgs(nx, ny, nat.dtype_code(self.dtype), S, el.lamda, el.mu, el.gamma,
                                       fields[0], fields[1], x0, x1, y0, y1, nat.ptr(ab), None, nat.ptr(out),
                                       nat.ptr(rbar), None, None, None,
                                       nat.ptr(acc["lam"]) if grads else None, nat.ptr(acc["mu"]) if grads else None,
                                       table.ref() if table is not None else None)
            ptrs = (ctypes.c_void_p * step)(*[nat.ptr(r) for r in recs]) if grads else None
            nat.check(nat.lib().xb_elast_smooth_adjoint_tiles(a, step, float(self.nlf.scale),
                                                              gt.ref() if gt is not None else None, ptrs,
                                                              nat.ptr(base), nat.stream()),
                      "xb_elast_smooth_adjoint_tiles")
            ab = out
            j -= step
        return ab

    def _level_adjoint(self, level, ab, states, rbar, acc):
        """Adjoint of one V-cycle from ``level`` down: ``ab`` is the adjoint of its
        output, ``states`` the finest-level sweep inputs (level 0) and ``rbar`` the
        accumulator of the level's residual adjoint (below level 0).  Returns the
        adjoint of the level's input populations."""
        el, k = self.el, self.el.k_smooth
        S, R = nat.XB_ADJ_SMOOTH, nat.XB_ADJ_RESIDUAL
        nx, ny = el.grid_size(level)
        small = states is None and level > 0 and self._small(level)
        post = states[k + 1:] if states is not None else [None] * k
        if small:
            ab = self._smooth_adjoint_small(level, ab, k, rbar)
        else:
            ab = self._smooth_adjoint_run(level, ab, post, rbar, acc)
        if level < el.levels - 1:
            nx, ny = el.grid_size(level)
            nxc, nyc = el.grid_size(level + 1)
            x0, x1, y0, y1 = el.interior_box(level)
            graph = level == 0 and getattr(self, "coarse_graph", None) is not None
            ab_c = self.ab_c if graph else torch.empty((8, nxc, nyc), dtype=self.dtype, device=self.device)
            nat.check(nat.lib().xb_elast_restrict_sum(nx, ny, nat.dtype_code(self.dtype), nat.ptr(ab), nat.ptr(ab_c),
                                                      nxc, nyc, x0, x1, y0, y1, nat.stream()), "xb_elast_restrict_sum")
            if graph:
                # rbar_c is read by the residual sweep below before the next replay
                self.coarse_graph.replay()
                rbar_c = self.rbar_c
            else:
                rbar_c = nat.fill(torch.empty((8, (x1 - x0) // 2, (y1 - y0) // 2), dtype=self.dtype,
                                              device=self.device), 0.0)
                self._level_adjoint(level + 1, ab_c, None, rbar_c, acc)
            ab = self._sweep_adjoint(level, R, rbar_c, ab, states[k] if states is not None else None, rbar, acc)
        pre = states[:k] if states is not None else [None] * k
        if small:
            return self._smooth_adjoint_small(level, ab, k, rbar)
        return self._smooth_adjoint_run(level, ab, pre, rbar, acc)

    def _small(self, level):
        """Whether ``level``'s transposed smoothing runs as one single-workgroup launch
        per run of sweeps: when its adjoint state, force adjoint and residual adjoint
        fit the kernel's 64 KB of LDS (18 values per cell and 8 per interior cell), or
        by the ``adjoint_small_cells`` threshold when one is set."""
        el = self.el
        nx, ny = el.grid_size(level)
        if self.adjoint_small_cells is not None:
            return nx * ny <= self.adjoint_small_cells
        x0, x1, y0, y1 = el.interior_box(level)
        size = torch.tensor([], dtype=self.dtype).element_size()
        return (18 * nx * ny + 8 * (x1 - x0) * (y1 - y0)) * size <= 65536

    def _smooth_adjoint_small(self, level, ab_in, sweeps, rbar):
        """``sweeps`` transposed smoothing sweeps of a coarse level (no recorded
        states) with their transposed coupling force in one launch
        (``xb_elast_smooth_adjoint_small``)."""
        el = self.el
        nx, ny = el.grid_size(level)
        x0, x1, y0, y1 = el.interior_box(level)
        out = torch.empty((8, nx, ny), dtype=self.dtype, device=self.device)
        tmp = torch.empty_like(out)
        fbar = torch.empty((2, nx, ny), dtype=self.dtype, device=self.device)
        table = el._tables.get(level)
        a = nat.ElastMGAdjointArgs(nx, ny, nat.dtype_code(self.dtype), nat.XB_ADJ_SMOOTH, el.lamda, el.mu, el.gamma,
                                   None, None, x0, x1, y0, y1, nat.ptr(ab_in), None, nat.ptr(out), nat.ptr(rbar),
                                   nat.ptr(fbar), None, None, None, None,
                                   table.ref() if table is not None else None)
        gt = self.da._grad_table(nx, ny)
        nat.check(nat.lib().xb_elast_smooth_adjoint_small(a, int(sweeps), float(self.nlf.scale),
                                                          gt.ref() if gt is not None else None, nat.ptr(tmp),
                                                          nat.stream()), "xb_elast_smooth_adjoint_small")
        return out

    def _inner_adjoint(self, t, fe_prev, base, ab, acc, last_recs):
        """Adjoint of the V-cycles of physical step ``t`` (``ab``: adjoint of their
        output); ``last_recs`` are the records of the last segment (already
        recomputed).  Returns the adjoint of the V-cycles' input."""
        starts = [fe_prev] + self.ckpt_cycle[t]
        for s in range(len(starts) - 1, -1, -1):
            if s == len(starts) - 1:
                recs = last_recs
            else:
                recs, _ = self._recompute(starts[s], self.K, base, self.fc[t][s * self.K:(s + 1) * self.K])
            for states in reversed(recs):
                ab = self._level_adjoint(0, ab, states, None, acc)
        return ab

    def _recompute_segment(self, t0, t1):
        """Physical steps t0 .. t1-1 recomputed from the checkpoint of step t0: per
        step its inputs, the coupler's pre-inner context, the records of its last
        V-cycle segment and the V-cycles' output (a list, last step at the end)."""
        p = self.p
        fe, fd, dde = self.ckpt_step[t0 // self.P]
        out = []
        for t in range(t0, t1):
            ctx = p._pre_inner(fe, fd, dde, t)
            starts = [fe] + self.ckpt_cycle[t]
            m_last = self.n - (len(starts) - 1) * self.K
            recs, fe_new = self._recompute(starts[-1], m_last, ctx["force"], self.fc[t][self.n - m_last:])
            out.append((fe, fd, dde, ctx, recs, fe_new))
            if t + 1 < t1:
                o = p._post_inner(ctx, fe_new)
                fe, fd, dde = o[0], o[1], o[4]
        return out

    # ---- transposed coupler
    def backward(self, eta_bar, p_bar, want_lame, want_omega):
        p, el, da = self.p, self.el, self.da
        T, dev = self.dtype, self.device
        nx, ny = el.nx, el.ny
        N = nx * ny
        dt = nat.dtype_code(T)
        th, al, eps = p.couplingRatio, p.alpha, p.epsilon
        lib = nat.lib()
        acc = {"lam": _zeros((1, nx, ny), torch.float64, dev) if want_lame else None,
               "mu": _zeros((1, nx, ny), torch.float64, dev) if want_lame else None,
               "omega": _zeros((1, nx, ny), torch.float64, dev)}
        g_l, g_m = nat.ptr(acc["lam"]), nat.ptr(acc["mu"])
        lf, mf = el._fields()
        gt0 = da._grad_table(nx, ny)
        gt0r = gt0.ref() if gt0 is not None else None
        ab_fe = _zeros((8, nx, ny), T, dev)
        ab_fd = _zeros((9, nx, ny), T, dev)
        ab_dde = _zeros((1, nx, ny), T, dev)
        eta_bar = None if eta_bar is None else eta_bar.to(T).contiguous()
        p_bar = None if p_bar is None else p_bar.to(T).contiguous()

        def lincomb(a, x, b, y, out):
            nat.check(lib.xb_lincomb(out.numel(), dt, float(a), nat.ptr(x), float(b), nat.ptr(y), nat.ptr(out),
                                     nat.stream()), "xb_lincomb")

        for t in range(self.steps - 1, -1, -1):
            if t == self.steps - 1 or (t + 1) % self.P == 0:
                segment = self._recompute_segment(t - t % self.P, t + 1)
            fe_prev, fd, dde, ctx, last_recs, fe_new = segment.pop()
            last = t == self.steps - 1
            base = ctx["force"]
            source = self._lincomb_new(ctx["source_base"], -(th * al / eps), el._divergence_soa(fe_new, fe_prev))
            acc["base"] = _zeros((2, nx, ny), torch.float64, dev)
            # Darcy step (poroelasticity.py:339)
            ab_fd_in = torch.empty_like(ab_fd)
            s64 = _zeros((1, nx, ny), torch.float64, dev)
            a = nat.DarcyAdjointArgs(nx, ny, dt, 0, da.omega, nat.ptr(da._omega_field), nat.ptr(da._source_soa),
                                     nat.ptr(source), nat.ptr(ab_fd), nat.ptr(fd), nat.ptr(ab_fd_in), None,
                                     nat.ptr(acc["omega"]), nat.ptr(s64),
                                     da._bc_table.ref() if da._bc_table is not None else None)
            nat.check(lib.xb_darcy_adjoint(a, nat.stream()), "xb_darcy_adjoint")
            s_bar = s64.to(T)
            # pressure = sum fd + (static + source) / 2
            if last and p_bar is not None:
                nat.check(lib.xb_darcy_macro_adjoint(nx, ny, dt, da.omega, nat.ptr(da._omega_field), nat.ptr(fd),
                                                     nat.ptr(p_bar), None, nat.ptr(ab_fd_in), nat.ptr(s_bar), None,
                                                     nat.stream()), "xb_darcy_macro_adjoint")
            # source = source_base - (theta alpha / eps) dde_new
            sb_bar = s_bar.clone()
            dde_bar = torch.empty_like(ab_dde)
            lincomb(1.0, ab_dde, -(th * al / eps), s_bar, dde_bar)
            # dde_new = div(fe_new) - div(fe_prev)
            ab_prev = _zeros((8, nx, ny), T, dev)
            nat.check(lib.xb_elast_divergence_adjoint(nx, ny, dt, el.lamda, el.mu, nat.ptr(fe_new), nat.ptr(fe_prev),
                                                      nat.ptr(dde_bar), nat.ptr(ab_fe), nat.ptr(ab_prev), lf, mf,
                                                      g_l, g_m, nat.stream()), "xb_elast_divergence_adjoint")
            # eta = m(fe_new) + F_body / 2 + ext / 2, ext = F0 + G_s(D fe_new)
            if last and eta_bar is not None:
                nat.check(lib.xb_elast_macro_adjoint(nx, ny, dt, el.lamda, el.mu, nat.ptr(fe_new), nat.ptr(eta_bar),
                                                     None, nat.ptr(ab_fe), lf, mf, None, None, nat.stream()),
                          "xb_elast_macro_adjoint")
                ext_bar = torch.empty((2, nx, ny), dtype=T, device=dev)
                lincomb(0.5, eta_bar, 0.0, None, ext_bar)
                nat.check(lib.xb_elast_coupling_adjoint(nx, ny, dt, el.lamda, el.mu, lf, mf, nat.ptr(ext_bar),
                                                        float(self.nlf.scale), gt0r, nat.ptr(ab_fe), nat.ptr(fe_new),
                                                        g_l, g_m, nat.ptr(acc["base"]), nat.stream()),
                          "xb_elast_coupling_adjoint")
            # the V-cycles of this step
            ab_in = self._inner_adjoint(t, fe_prev, base, ab_fe, acc, last_recs)
            lincomb(1.0, ab_prev, 1.0, ab_in, ab_prev)
            # F0 = -alpha eps pgb - G_s(D fe_prev)
            F0_bar = acc["base"].to(T)
            neg = torch.empty_like(F0_bar)
            lincomb(-1.0, F0_bar, 0.0, None, neg)
            nat.check(lib.xb_elast_coupling_adjoint(nx, ny, dt, el.lamda, el.mu, lf, mf, nat.ptr(neg),
                                                    float(self.nlf.scale), gt0r, nat.ptr(ab_prev), nat.ptr(fe_prev),
                                                    g_l, g_m, None, nat.stream()), "xb_elast_coupling_adjoint")
            pgb_bar = torch.empty_like(F0_bar)
            lincomb(-al * eps, F0_bar, 0.0, None, pgb_bar)
            if p.alternative_gradient:
                # pgb = G_1(p_base), p_base = sum fd + (static + source_base) / 2
                pb_bar = _zeros((1, nx, ny), T, dev)
                nat.check(lib.xb_darcy_gradient_adjoint(nx, ny, dt, nat.ptr(pgb_bar), 1.0, gt0r, 1.0, nat.ptr(pb_bar),
                                                        nat.stream()), "xb_darcy_gradient_adjoint")
                nat.check(lib.xb_darcy_macro_adjoint(nx, ny, dt, da.omega, nat.ptr(da._omega_field), nat.ptr(fd),
                                                     nat.ptr(pb_bar), None, nat.ptr(ab_fd_in), nat.ptr(sb_bar), None,
                                                     nat.stream()), "xb_darcy_macro_adjoint")
            else:
                # pgb = -3 omega sum c fd + G_1/2(source_base)
                nat.check(lib.xb_darcy_macro_adjoint(nx, ny, dt, da.omega, nat.ptr(da._omega_field), nat.ptr(fd),
                                                     None, nat.ptr(pgb_bar), nat.ptr(ab_fd_in), None,
                                                     nat.ptr(acc["omega"]), nat.stream()), "xb_darcy_macro_adjoint")
                nat.check(lib.xb_darcy_gradient_adjoint(nx, ny, dt, nat.ptr(pgb_bar), 0.5, gt0r, 1.0, nat.ptr(sb_bar),
                                                        nat.stream()), "xb_darcy_gradient_adjoint")
            # source_base = (1 - theta) (-alpha / eps) dde
            ab_dde_in = torch.empty_like(ab_dde)
            lincomb((1.0 - th) * -al / eps, sb_bar, 0.0, None, ab_dde_in)
            ab_fe, ab_fd, ab_dde = ab_prev, ab_fd_in, ab_dde_in
        return acc, (ab_fe, ab_fd, ab_dde)

    def _lincomb_new(self, x, b, y):
        out = torch.empty_like(x)
        nat.check(nat.lib().xb_lincomb(out.numel(), nat.dtype_code(x.dtype), 1.0, nat.ptr(x), float(b), nat.ptr(y),
                                       nat.ptr(out), nat.stream()), "xb_lincomb")
        return out


class _BiotSolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega_field, lam_field, mu_field, problem, steps, K):
        sessions = problem.__dict__.setdefault("_biot_adjoint_sessions", {})
        key = (int(steps), int(K))
        if key not in sessions:
            # the fields must be in place before the session captures its V-cycle graph
            el, da = problem.elasticity, problem.darcy
            el.set_lame_fields(lam_field.detach().reshape(el.nx, el.ny), mu_field.detach().reshape(el.nx, el.ny))
            da.set_omega_field(omega_field.detach().reshape(da.nx, da.ny))
            problem._inner = None  # a graph captured before the fields existed would not read them
            sessions[key] = BiotAdjointSession(problem, steps, K)
        s = sessions[key]
        s.bind(omega_field, lam_field, mu_field)
        eta, pressure = s.forward()
        ctx.s, ctx.generation = s, s.generation
        ctx.meta = [(t.shape, t.dtype) for t in (omega_field, lam_field, mu_field)]
        return eta, pressure

    @staticmethod
    def backward(ctx, eta_bar, p_bar):
        s = ctx.s
        if s.generation != ctx.generation:
            raise RuntimeError("the checkpoints of this coupled solve were overwritten by a later solve of the same "
                               "problem and shape; run its backward first")
        want_lame = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        acc, _ = s.backward(eta_bar, p_bar, want_lame, ctx.needs_input_grad[0])
        (so, do), (sl, dl), (sm, dm) = ctx.meta
        g_om = acc["omega"].reshape(so).to(do) if ctx.needs_input_grad[0] else None
        g_l = acc["lam"].reshape(sl).to(dl) if ctx.needs_input_grad[1] else None
        g_m = acc["mu"].reshape(sm).to(dm) if ctx.needs_input_grad[2] else None
        return g_om, g_l, g_m, None, None, None


def biot_solve(problem, omega_field, lam_field, mu_field, steps, checkpoint_every=50):
    """``(eta, pressure)`` of the last of ``steps`` coupled steps of ``problem`` (an
    :class:`LBMPoroelasticity` with multigrid elasticity and the fixed
    ``elasticity_steps`` V-cycles per step), SoA ``(2, nx, ny)`` and ``(1, nx, ny)``,
    with the per-cell Darcy rate ``omega_field`` and Lame fields ``lam_field``,
    ``mu_field`` (each ``(nx, ny)`` or ``(1, nx, ny)``).  Differentiable with respect
    to the three fields; V-cycle checkpoints every ``checkpoint_every`` V-cycles."""
    problem.elasticity.precisionPolicy.require_single("biot_solve")
    return _BiotSolveFn.apply(omega_field, lam_field, mu_field, problem, int(steps), int(checkpoint_every))
