"""Reverse-mode gradients of the pseudo-time-stepping solvers.

The reference differentiates nothing itself; the drivers' parameter studies would
use ``jax.grad`` through the jitted steps.  Here the gradient of a misfit of the
final populations with respect to the per-cell parameter fields comes from
hand-written adjoint step kernels (``xb_elast_adjoint`` / ``xb_darcy_adjoint``:
the transposed collide -> rules -> stream -> rules step, plus the parameter
derivatives of the step accumulated in FP64), wrapped in ``torch.autograd.Function``:

* :func:`elasticity_solve` - ``n`` plain D2Q8 steps with per-cell lambda(x), mu(x)
  (``LBMElasticity._step``, reference src/elasticity.py:1169-1219), differentiable
  with respect to lambda, mu, the external force and the initial populations;
* :func:`darcy_solve` - ``n`` D2Q9 steps with a per-cell relaxation rate omega(x)
  (``AdvectionDiffusionReactionBGK.step``, src/darcy.py:131-171), differentiable with
  respect to omega (hence the permeability, omega = 1 / (3 K + 1/2)), the source and
  the initial populations.

Backward is the discrete adjoint of exactly those ``n`` steps.  It recomputes the
states from checkpoints stored every ``checkpoint_every`` steps (one segment of
states is held at a time, so memory stays at ``n / K`` checkpoints plus ``K``
states), and replays the recompute and the adjoint steps of a segment as captured
HIP graphs.  The buffers and graphs of a solve are kept on the solver object per
(n, K, operands) and reused by later calls (an optimisation loop captures once);
a call's gradients are returned as new tensors.  A later solve of the same key
overwrites the checkpoints of an earlier one: the earlier one's backward then
raises (solve with ``reuse=False`` to keep several solves alive at once).  :func:`elasticity_steady_adjoint` / :func:`darcy_steady_adjoint` give
the gradient of a converged steady state by the fixed-point (implicit) adjoint:
iterate ``a <- (dS/df)^T a + dJ/df`` to convergence, then one parameter pass.
"""
import math

import torch

from . import _native as nat
from .base import ConvergedStepRunner, graph_capture


def _f64_zeros(shape, device):
    return torch.zeros(shape, dtype=torch.float64, device=device)


def _soa(x, c, nx, ny, dtype, device):
    """A SoA ``(c, nx, ny)`` tensor (``(nx, ny)`` accepted for c = 1) on the device."""
    if x is None:
        return None
    t = torch.as_tensor(x)
    if c == 1 and tuple(t.shape) == (nx, ny):
        t = t.reshape(1, nx, ny)
    if tuple(t.shape) != (c, nx, ny):
        raise ValueError(f"expected a ({c}, {nx}, {ny}) structure-of-arrays tensor, got {tuple(t.shape)}")
    return t.to(device=device, dtype=dtype).contiguous()


class ElasticityStepper:
    """Forward and adjoint single-grid steps of an :class:`LBMElasticity` with bound
    per-cell Lame fields (SoA ``(1, nx, ny)``) and an optional external force
    ``(2, nx, ny)``."""

    Q = 8

    def __init__(self, el, lam_field, mu_field, force=None):
        if el.levels != 1:
            raise ValueError("adjoint steps are single-grid (MG_levels = 1)")
        self.el = el
        # owned copies: captured graphs keep these addresses, bind() refreshes values
        self.lam = _soa(lam_field, 1, el.nx, el.ny, el._tdt, el.device).clone()
        self.mu = _soa(mu_field, 1, el.nx, el.ny, el._tdt, el.device).clone()
        self.force = None if force is None else _soa(force, 2, el.nx, el.ny, el._tdt, el.device).clone()
        self.shape = (8, el.nx, el.ny)
        self.dtype, self.device = el._tdt, el.device
        table = el._tables.get(0)
        self._bc = table.ref() if table is not None else None

    def bind(self, lam_field, mu_field, force=None):
        el = self.el
        self.lam.copy_(_soa(lam_field, 1, el.nx, el.ny, el._tdt, el.device))
        self.mu.copy_(_soa(mu_field, 1, el.nx, el.ny, el._tdt, el.device))
        if self.force is not None:
            self.force.copy_(_soa(force, 2, el.nx, el.ny, el._tdt, el.device))

    def step(self, src, dst):
        a = self.el._launch(src, dst, 0, nat.XB_MODE_STEP, ext=self.force)
        a.lam_field, a.mu_field = nat.ptr(self.lam), nat.ptr(self.mu)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step")
        return dst

    def adjoint(self, ab_in, f_in, ab_out, grads=None, ab_add=None):
        """``ab_out = (dS/df)^T ab_in (+ ab_add)``; with ``grads`` (dict of FP64
        accumulators "lam", "mu", "force"), add the parameter derivatives at ``f_in``."""
        g = grads or {}
        el = self.el
        a = nat.ElastAdjointArgs(el.nx, el.ny, nat.dtype_code(self.dtype), 0, el.lamda, el.mu,
                                 nat.ptr(self.lam), nat.ptr(self.mu), nat.ptr(ab_in), nat.ptr(f_in), nat.ptr(ab_out),
                                 nat.ptr(ab_add), nat.ptr(g.get("lam")), nat.ptr(g.get("mu")),
                                 nat.ptr(g.get("force")), self._bc)
        nat.check(nat.lib().xb_elast_adjoint(a, nat.stream()), "xb_elast_adjoint")
        return ab_out

    def grad_buffers(self, want_force):
        n = (1, self.el.nx, self.el.ny)
        g = {"lam": _f64_zeros(n, self.device), "mu": _f64_zeros(n, self.device)}
        if want_force:
            g["force"] = _f64_zeros((2, self.el.nx, self.el.ny), self.device)
        return g


class DarcyStepper:
    """Forward and adjoint steps of an :class:`AdvectionDiffusionReactionBGK` with a
    bound per-cell relaxation rate (SoA ``(1, nx, ny)``) and optional source."""

    Q = 9

    def __init__(self, da, omega_field, source=None):
        self.da = da
        self.omega = _soa(omega_field, 1, da.nx, da.ny, da._tdt, da.device).clone()
        self.source = None if source is None else _soa(source, 1, da.nx, da.ny, da._tdt, da.device).clone()
        self.shape = (9, da.nx, da.ny)
        self.dtype, self.device = da._tdt, da.device
        self._bc = da._bc_table.ref() if da._bc_table is not None else None

    def bind(self, omega_field, source=None):
        da = self.da
        self.omega.copy_(_soa(omega_field, 1, da.nx, da.ny, da._tdt, da.device))
        if self.source is not None:
            self.source.copy_(_soa(source, 1, da.nx, da.ny, da._tdt, da.device))

    def step(self, src, dst):
        a = self.da._args(src, dst, self.source, None)
        a.omega_field = nat.ptr(self.omega)
        nat.check(nat.lib().xb_darcy_step(a, nat.stream()), "xb_darcy_step")
        return dst

    def adjoint(self, ab_in, f_in, ab_out, grads=None, ab_add=None):
        g = grads or {}
        da = self.da
        a = nat.DarcyAdjointArgs(da.nx, da.ny, nat.dtype_code(self.dtype), 0, da.omega, nat.ptr(self.omega),
                                 nat.ptr(da._source_soa), nat.ptr(self.source), nat.ptr(ab_in), nat.ptr(f_in),
                                 nat.ptr(ab_out), nat.ptr(ab_add), nat.ptr(g.get("omega")), nat.ptr(g.get("source")),
                                 self._bc)
        nat.check(nat.lib().xb_darcy_adjoint(a, nat.stream()), "xb_darcy_adjoint")
        return ab_out

    def grad_buffers(self, want_source):
        n = (1, self.da.nx, self.da.ny)
        g = {"omega": _f64_zeros(n, self.device)}
        if want_source:
            g["source"] = _f64_zeros(n, self.device)
        return g


class CheckpointedSolve:
    """``n`` steps of a stepper with checkpoints every ``K`` steps, and the backward
    pass: per segment (last first), recompute its states from the checkpoint into a
    segment buffer, then run the adjoint steps through them, accumulating the
    parameter gradients.  Full segments replay captured graphs."""

    def __init__(self, stepper, n, K, use_graphs=True):
        if n < 1 or K < 1:
            raise ValueError("need n >= 1 steps and a checkpoint interval K >= 1")
        self.s, self.n, self.K = stepper, n, min(K, n)
        self.nseg = math.ceil(n / self.K)
        self.use_graphs = use_graphs and torch.cuda.is_available()
        shape, dt, dev = stepper.shape, stepper.dtype, stepper.device
        self.ckpt = torch.empty((self.nseg,) + shape, dtype=dt, device=dev)
        self.states = torch.empty((self.K,) + shape, dtype=dt, device=dev)
        self.ab = [torch.empty(shape, dtype=dt, device=dev) for _ in range(2)]
        self._fwd_graph = self._rec_graph = self._adj_graph = None
        self._grads = None
        self.grads = None  # FP64 accumulators owned by the session (see gradient_buffers)
        self._grads_extra = None
        self.generation = 0  # forward passes run; backward checks it saw the last one

    def gradient_buffers(self, want_extra):
        """The session's FP64 accumulators, zeroed (kept across calls so that the
        captured backward graphs stay valid)."""
        if self.grads is None or self._grads_extra != bool(want_extra):
            self.grads = self.s.grad_buffers(want_extra)
            self._grads_extra = bool(want_extra)
        for g in self.grads.values():
            g.zero_()
        return self.grads

    # forward
    def forward(self, f0):
        self.generation += 1
        s, K = self.s, self.K
        a = self.states[0]
        b = self.states[1] if K > 1 else torch.empty_like(a)
        done = 0
        for j in range(self.nseg):
            self.ckpt[j].copy_(f0 if j == 0 else a)
            if j == 0:
                a.copy_(f0)
            m = min(K, self.n - done)
            if self.use_graphs and m == K and K % 2 == 0 and K >= 2:
                if self._fwd_graph is None:
                    keep = a.clone()
                    self._fwd_graph = self._capture(lambda: self._pingpong(a, b, K))
                    a.copy_(keep)  # the eager warm-up advanced the state
                self._fwd_graph.replay()
            else:
                self._pingpong(a, b, m)
                if m % 2:
                    a.copy_(b)
            done += m
        return a.clone()

    def _pingpong(self, a, b, m):
        for t in range(m):
            if t % 2 == 0:
                self.s.step(a, b)
            else:
                self.s.step(b, a)

    @staticmethod
    def _capture(fn):
        fn()  # eager warm-up outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with graph_capture(g):
            fn()
        return g

    # backward
    def _recompute(self, m):
        for t in range(1, m):
            self.s.step(self.states[t - 1], self.states[t])

    def _adjoint_segment(self, m, grads):
        # ab[0] holds the adjoint of the segment's last output; steps t = m-1 .. 0
        for i, t in enumerate(range(m - 1, -1, -1)):
            src, dst = self.ab[i % 2], self.ab[(i + 1) % 2]
            self.s.adjoint(src, self.states[t], dst, grads)
        if m % 2:
            self.ab[0].copy_(self.ab[1])

    def backward(self, grad_out, grads):
        """Adjoint of the initial populations, given that of the final ones; adds the
        parameter gradients to ``grads`` (FP64 buffers from the stepper)."""
        self.ab[0].copy_(grad_out)
        K = self.K
        for j in range(self.nseg - 1, -1, -1):
            m = min(K, self.n - j * K)
            self.states[0].copy_(self.ckpt[j])
            full = self.use_graphs and m == K and K % 2 == 0 and K >= 2
            if full:
                if self._rec_graph is None or self._grads is not grads:
                    # the eager warm-ups of the captures run the segment once: keep the
                    # accumulators and the incoming adjoint (states[0] is not written)
                    saved = [g.clone() for g in grads.values()], self.ab[0].clone()
                    self._rec_graph = self._capture(lambda: self._recompute(K))
                    self._adj_graph = self._capture(lambda: self._adjoint_segment(K, grads))
                    for g, v in zip(grads.values(), saved[0]):
                        g.copy_(v)
                    self.ab[0].copy_(saved[1])
                    self._grads = grads
                self._rec_graph.replay()
                self._adj_graph.replay()
            else:
                self._recompute(m)
                self._adjoint_segment(m, grads)
        return self.ab[0].clone()


def _check_generation(ctx):
    if ctx.cs.generation != ctx.generation:
        raise RuntimeError("the checkpoints of this solve were overwritten by a later solve of the same solver, "
                           "steps and interval; run its backward first, or solve with reuse=False")


class _ElasticitySolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lam_field, mu_field, force, f0, el, n, K, use_graphs, reuse):
        key = ("elasticity", n, K, use_graphs, force is not None)
        sessions = el.__dict__.setdefault("_adjoint_sessions", {}) if reuse else {}
        if key in sessions:
            cs = sessions[key]
            st = cs.s
            st.bind(lam_field.detach(), mu_field.detach(), None if force is None else force.detach())
        else:
            st = ElasticityStepper(el, lam_field.detach(), mu_field.detach(),
                                   None if force is None else force.detach())
            cs = sessions[key] = CheckpointedSolve(st, n, K, use_graphs)
        out = cs.forward(_soa(f0.detach(), 8, el.nx, el.ny, el._tdt, el.device))
        ctx.cs, ctx.st, ctx.generation = cs, st, cs.generation
        ctx.has_force = force is not None
        ctx.shapes = (lam_field.shape, mu_field.shape, None if force is None else force.shape, f0.shape)
        ctx.dtypes = (lam_field.dtype, mu_field.dtype, None if force is None else force.dtype, f0.dtype)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        st, cs = ctx.st, ctx.cs
        _check_generation(ctx)
        need_force = ctx.has_force and ctx.needs_input_grad[2]
        grads = cs.gradient_buffers(need_force)
        ab0 = cs.backward(grad_out.contiguous().to(st.dtype), grads)
        sl, sm, sf, s0 = ctx.shapes
        dl, dm, df, d0 = ctx.dtypes
        g_lam = grads["lam"].reshape(sl).to(dl) if ctx.needs_input_grad[0] else None
        g_mu = grads["mu"].reshape(sm).to(dm) if ctx.needs_input_grad[1] else None
        g_f = grads["force"].reshape(sf).to(df) if need_force else None
        g_0 = ab0.reshape(s0).to(d0) if ctx.needs_input_grad[3] else None
        return g_lam, g_mu, g_f, g_0, None, None, None, None, None


def elasticity_solve(el, lam_field, mu_field, f0, n, force=None, checkpoint_every=50, use_graphs=True, reuse=True):
    """Populations after ``n`` single-grid steps of ``el`` (SoA ``(8, nx, ny)``) with
    per-cell Lame fields ``(1, nx, ny)`` (or ``(nx, ny)``), an optional external force
    ``(2, nx, ny)`` and initial populations ``f0`` ``(8, nx, ny)``, all SoA.  Differentiable with respect
    to the fields, the force and ``f0``."""
    el.precisionPolicy.require_single("elasticity_solve")
    return _ElasticitySolveFn.apply(lam_field, mu_field, force, f0, el, int(n), int(checkpoint_every), use_graphs,
                                    reuse)


class _DarcySolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega_field, source, f0, da, n, K, use_graphs, reuse):
        key = ("darcy", n, K, use_graphs, source is not None)
        sessions = da.__dict__.setdefault("_adjoint_sessions", {}) if reuse else {}
        if key in sessions:
            cs = sessions[key]
            st = cs.s
            st.bind(omega_field.detach(), None if source is None else source.detach())
        else:
            st = DarcyStepper(da, omega_field.detach(), None if source is None else source.detach())
            cs = sessions[key] = CheckpointedSolve(st, n, K, use_graphs)
        out = cs.forward(_soa(f0.detach(), 9, da.nx, da.ny, da._tdt, da.device))
        ctx.cs, ctx.st, ctx.generation = cs, st, cs.generation
        ctx.has_source = source is not None
        ctx.shapes = (omega_field.shape, None if source is None else source.shape, f0.shape)
        ctx.dtypes = (omega_field.dtype, None if source is None else source.dtype, f0.dtype)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        st, cs = ctx.st, ctx.cs
        _check_generation(ctx)
        need_src = ctx.has_source and ctx.needs_input_grad[1]
        grads = cs.gradient_buffers(need_src)
        ab0 = cs.backward(grad_out.contiguous().to(st.dtype), grads)
        so, ss, s0 = ctx.shapes
        do, ds, d0 = ctx.dtypes
        g_om = grads["omega"].reshape(so).to(do) if ctx.needs_input_grad[0] else None
        g_s = grads["source"].reshape(ss).to(ds) if need_src else None
        g_0 = ab0.reshape(s0).to(d0) if ctx.needs_input_grad[2] else None
        return g_om, g_s, g_0, None, None, None, None, None


def darcy_solve(da, omega_field, f0, n, source=None, checkpoint_every=50, use_graphs=True, reuse=True):
    """Populations after ``n`` steps of ``da`` (SoA ``(9, nx, ny)``) with a per-cell
    relaxation rate ``omega_field`` and an optional external source ``(1, nx, ny)``.
    Differentiable with respect to omega, the source and ``f0``.  For the
    permeability, omega = 1 / (3 K + 1/2): pass ``1 / (3 * K + 0.5)`` built from a
    ``K`` that requires grad."""
    da.precisionPolicy.require_single("darcy_solve")
    return _DarcySolveFn.apply(omega_field, source, f0, da, int(n), int(checkpoint_every), use_graphs, reuse)


class _AdjointIteration:
    """The iteration ``a <- (dS/df)^T a + b`` of one stepper at a fixed state, on the
    device-tested step loop (:class:`ConvergedStepRunner`: captured blocks of
    ``check_every`` adjoint steps, the relative step-to-step norm latched on the
    device).  The fixed state and ``b`` live in buffers of its own, refreshed per call."""

    def __init__(self, stepper, check_every, checks):
        self.f = torch.zeros(stepper.shape, dtype=stepper.dtype, device=stepper.device)
        self.b = torch.zeros(stepper.shape, dtype=stepper.dtype, device=stepper.device)
        f, b = self.f, self.b
        self.runner = ConvergedStepRunner(lambda src, dst: stepper.adjoint(src, f, dst, None, b), self.b,
                                          check_every, checks)


def _fixed_point_adjoint(stepper, f_star, dJdf, tol, max_steps, check_every, want_extra, checks=10):
    """Solve ``a = (dS/df)^T a + dJ/df`` by iteration from ``a_0 = dJ/df`` until
    ``|a_k - a_{k-1}| <= tol |a_1 - a_0|``, then add ``a . dS/dtheta`` at the fixed
    point ``f_star`` to fresh FP64 gradient buffers.  Returns ``(grads, a, k, rel)``;
    ``k`` is ``max_steps`` when the iteration did not get there."""
    key = (int(check_every), int(checks))
    cache = stepper.__dict__.setdefault("_adjoint_iterations", {})
    it = cache.get(key)
    if it is None:
        it = cache[key] = _AdjointIteration(stepper, check_every, checks)
    it.f.copy_(f_star)
    it.b.copy_(dJdf.contiguous().to(stepper.dtype))
    a, k, rel = it.runner.run(it.b, tol, max_steps)
    if not math.isfinite(rel):
        raise FloatingPointError(f"adjoint iteration is not finite after {k} steps")
    if k < 0:
        k = int(max_steps)
    grads = stepper.grad_buffers(want_extra)
    scratch = torch.empty_like(a)
    stepper.adjoint(a, it.f, scratch, grads)
    return grads, a, k, rel


def _steady_stepper(solver, key, make, bind):
    """The solver's stepper for steady adjoints of this kind, kept with its captured
    adjoint iterations and rebound to the call's parameters."""
    cache = solver.__dict__.setdefault("_steady_steppers", {})
    st = cache.get(key)
    if st is None:
        st = cache[key] = make()
    else:
        bind(st)
    return st


def elasticity_steady_adjoint(el, lam_field, mu_field, f_star, dJdf, force=None, tol=1e-10, max_steps=200000,
                              check_every=50):
    """Gradient of ``J(f*)`` for the steady state ``f* = S(f*)`` of the single-grid
    elasticity step with per-cell Lame fields: returns ``(grads, a, steps, rel)``
    with ``grads["lam"]``, ``grads["mu"]`` (and ``grads["force"]``) FP64 fields."""
    el.precisionPolicy.require_single("elasticity_steady_adjoint")
    st = _steady_stepper(el, ("elasticity", force is not None),
                         lambda: ElasticityStepper(el, lam_field, mu_field, force),
                         lambda st: st.bind(lam_field, mu_field, force))
    f = _soa(f_star, 8, el.nx, el.ny, el._tdt, el.device)
    return _fixed_point_adjoint(st, f, dJdf, tol, max_steps, check_every, force is not None)


def darcy_steady_adjoint(da, omega_field, f_star, dJdf, source=None, tol=1e-10, max_steps=200000, check_every=50):
    """As :func:`elasticity_steady_adjoint` for the Darcy step: ``grads["omega"]``
    (and ``grads["source"]``)."""
    da.precisionPolicy.require_single("darcy_steady_adjoint")
    st = _steady_stepper(da, ("darcy", source is not None), lambda: DarcyStepper(da, omega_field, source),
                         lambda st: st.bind(omega_field, source))
    f = _soa(f_star, 9, da.nx, da.ny, da._tdt, da.device)
    return _fixed_point_adjoint(st, f, dJdf, tol, max_steps, check_every, source is not None)
