"""D2Q9 BGK advection-diffusion-reaction solver used for Darcy (pressure-diffusion) flow.

Drop-in for the reference ``src/darcy.py`` (``AdvectionDiffusionReactionBGK``
:22-296 and the ADR boundary conditions :298-466).  The step, the macroscopic
update and the lattice gradient run on the GPU through ``include/xlb_biot.h``.

Extension: ``omega`` may also be a per-cell field (shape ``(nx, ny)`` or
``(nx, ny, 1)``) for heterogeneous permeability ``K = (1/omega - 1/2)/3``; a float
reproduces the reference exactly.  Advection (``vel``) is not on this path.
"""
import time
import weakref

import numpy as np
import torch

from . import _native as nat
from . import autotune
from .base import ConvergedStepRunner, LBMBase, StepRunner, as_soa, io_field, view
from .boundary_conditions import BoundaryCondition, BounceBackHalfway, RuleTable, build_rule_table, compile_rules


class AdvectionDiffusionReactionBGK(LBMBase):
    """reference darcy.py:22-34."""

    def __init__(self, **kwargs):
        omega = kwargs.get("omega")
        self._omega_field = None
        if omega is not None and not isinstance(omega, float):
            # per-cell relaxation rate; the scalar slot keeps its mean for reporting
            field = np.asarray(omega.detach().cpu() if isinstance(omega, torch.Tensor) else omega, dtype=np.float64)
            kwargs = dict(kwargs)
            kwargs["omega"] = float(field.mean())
            self._omega_field_host = field
        else:
            self._omega_field_host = None
        super().__init__(**kwargs)
        if self._omega_field_host is not None:
            self._omega_field = as_soa(self._omega_field_host.reshape(self.nx, self.ny, 1), self.nx, self.ny, 1,
                                       self._tdt, self.device)
        self.vel = kwargs.get("vel", None)
        if self.vel is not None:
            raise NotImplementedError("advection (vel) is not supported on this path")
        src = self.get_source()
        self._source_soa = None if src is None else as_soa(src, self.nx, self.ny, 1, self._tdt, self.device)
        self.source = None if self._source_soa is None else view(self._source_soa)
        self._grad_tables = {}
        self._runners = {}
        self.mlups = None

    @property
    def _tdt(self):
        return self.precisionPolicy.torch_compute

    def _code(self, t):
        """Kernel dtype of a population buffer: the storage-rounding dtype under the
        ``"f64/f32"``, ``"f64/f16"`` and ``"f32/f16"`` policies."""
        return self.precisionPolicy.storage_code if self.precisionPolicy.mixed else nat.dtype_code(t.dtype)

    def _out(self, soa):
        """A population state handed to the caller, in the storage precision."""
        if soa is None:
            return None
        if self.precisionPolicy.mixed:
            soa = soa.to(self.precisionPolicy.torch_output)
        return view(soa)

    @property
    def omega_field(self):
        """Per-cell relaxation rate ``(nx, ny, 1)`` or None."""
        return None if self._omega_field is None else view(self._omega_field)

    def set_omega_field(self, omega_field):
        """Per-cell relaxation rate ``(nx, ny[, 1])`` (copied; a later call overwrites
        the field in place, so captured graphs keep reading it)."""
        field = as_soa(omega_field, self.nx, self.ny, 1, self._tdt, self.device)
        if self._omega_field is not None:
            self._omega_field.copy_(field)
            return
        self._omega_field = field.clone()
        self._runners = {}

    def _soa(self, x, c, nx=None, ny=None):
        if x is None:
            return None
        return as_soa(x, nx or self.nx, ny or self.ny, c, self._tdt, self.device)

    def _new(self, c, nx=None, ny=None):
        return torch.empty((c, nx or self.nx, ny or self.ny), dtype=self._tdt, device=self.device)

    # ---- darcy.py:36-73
    def _macro_soa(self, f, ext_source=None, p_out=True, grad_out=False):
        p = self._new(1) if p_out else None
        g = self._new(2) if grad_out else None
        nat.check(nat.lib().xb_darcy_macro(self.nx, self.ny, nat.dtype_code(self._tdt), self.omega,
                                           nat.ptr(self._omega_field), nat.ptr(f), nat.ptr(self._source_soa),
                                           nat.ptr(ext_source), nat.ptr(p), nat.ptr(g), nat.stream()),
                  "xb_darcy_macro")
        return p, g

    def update_macroscopic(self, f, external_source=None, gradient=False):
        fs = self._soa(f, 9)
        p, g = self._macro_soa(fs, self._soa(external_source, 1), True, gradient)
        if gradient:
            return view(p), view(g)
        return view(p)

    # ---- darcy.py:75-104
    def _grad_table(self, nx, ny):
        """Rule table of the bounce-back / anti-bounce-back walls for a ``(nx, ny)`` scalar.

        On another shape (coarse multigrid levels, reached through the Biot
        nonlinear force) the fine-grid wall indices are used as they are and rows
        outside the array are dropped, as JAX's ``.at[].set`` does."""
        key = (nx, ny)
        if key not in self._grad_tables:
            bcs = [bc for bc in self.BCs if bc.name in ("ADR_Neumann_BounceBack", "ADR_AntiBounceBack")]
            if key == (self.nx, self.ny):
                t = build_rule_table(bcs, nx, ny, self.q, self._tdt, self.device, stages=("PostStreaming",))
            else:
                cells, rules = compile_rules(bcs, self.nx, self.ny, self.q, stages=("PostStreaming",),
                                             bounds=(nx, ny))
                t = None
                if len(cells):
                    code, const = rules["PostStreaming"]
                    t = RuleTable(nx, ny, self.q, (cells // self.ny) * ny + cells % self.ny, None, None, code,
                                  const, self._tdt, self.device)
            self._grad_tables[key] = t
        return self._grad_tables[key]

    def _gradient_soa(self, q, scale=1.0, base=None, a=1.0, b=1.0, out=None):
        nx, ny = q.shape[1], q.shape[2]
        table = self._grad_table(nx, ny)
        if out is None:
            out = torch.empty((2, nx, ny), dtype=q.dtype, device=q.device)
        nat.check(nat.lib().xb_darcy_gradient(nx, ny, nat.dtype_code(q.dtype), nat.ptr(q), float(scale),
                                              table.ref() if table is not None else None, nat.ptr(base),
                                              float(a), float(b), nat.ptr(out), nat.stream()),
                  "xb_darcy_gradient")
        return out

    def gradient(self, q, scale=1.0):
        """Lattice gradient ``3 scale sum_i w_i c_i q(x + c_i)`` with BB/ABB walls."""
        if isinstance(q, torch.Tensor) and q.dim() == 3:
            nx, ny = q.shape[0], q.shape[1]
        else:
            nx, ny = self.nx, self.ny
        return view(self._gradient_soa(self._soa(q, 1, nx, ny), scale))

    # ---- darcy.py:106-171
    def _tune_key(self, two=False):
        return (f"darcy/{self.precision}/{self.nx}x{self.ny}/bc{int(self._bc_table is not None)}"
                f"/omf{int(self._omega_field is not None)}" + ("/x2" if two else ""))

    def _args(self, f, out, ext_source, fpc, two=False):
        table = self._bc_table
        return nat.DarcyArgs(self.nx, self.ny, self._code(f), 0, self.omega, nat.ptr(self._omega_field),
                             nat.ptr(f), nat.ptr(out), nat.ptr(fpc), nat.ptr(self._source_soa), nat.ptr(ext_source),
                             table.ref() if table is not None else None, autotune.lookup(self._tune_key(two)), 0)

    def _step2_into(self, f, out, ext_source=None):
        """Two steps in one pass (``xb_darcy_step2``)."""
        nat.check(nat.lib().xb_darcy_step2(self._args(f, out, ext_source, None, True), nat.stream()),
                  "xb_darcy_step2")
        return out

    def autotune(self):
        """Measure and cache the best sweep launch shape of this solver's step, then
        the best number of steps per launch for step loops: one or two per sweep, or
        2 .. 8 on LDS tiles (``xb_darcy_stepn``)."""
        lib = nat.lib()
        f = torch.zeros((9, self.nx, self.ny), dtype=self._tdt, device=self.device)
        o = torch.empty_like(f)
        a = self._args(f, o, None, None)

        def run(rows):
            a.rows = rows
            nat.check(lib.xb_darcy_step(a, nat.stream()), "xb_darcy_step(tune)")
        key = self._tune_key()
        out = {key: autotune.tune(key, run)}
        launches = {autotune.SWEEP1: lambda: self._step_soa(f, None, False, out=o)}
        if min(self.nx, self.ny) >= 2:
            a2 = self._args(f, o, None, None, True)

            def run2(rows):
                a2.rows = rows
                nat.check(lib.xb_darcy_step2(a2, nat.stream()), "xb_darcy_step2(tune)")
            key2 = self._tune_key(True)
            out[key2] = autotune.tune(key2, run2)
            launches[autotune.SWEEP2] = lambda: self._step2_into(f, o)
        for m in autotune.TILE_STEPS:
            for nt in autotune.TILE_THREADS:
                launches[autotune.tile_choice(m, nt)] = (lambda m=m, nt=nt: self._stepn_into(f, o, m, threads=nt))
        km = self._tune_key() + "/multi"
        out[km] = autotune.tune_multi(km, launches)
        return out

    def _stepn_into(self, f, out, m, ext_source=None, threads=0):
        """``m`` steps in one launch on LDS tiles (``xb_darcy_stepn``) of ``threads``
        threads per block (0: the library's choice)."""
        a = self._args(f, out, ext_source, None)
        a.rows = int(threads)
        nat.check(nat.lib().xb_darcy_stepn(a, int(m), nat.stream()), "xb_darcy_stepn")
        return out

    def _multi(self):
        """The plain-step launch of step loops (choice, steps per launch,
        launch(src, dst)): ``self.step_launch`` when set (an ``autotune`` choice),
        else the tuned one.  Narrower storage (``"f64/f32"``, ``"f64/f16"``,
        ``"f32/f16"``) is rounded by the single-step sweep only."""
        if self.precisionPolicy.mixed:
            return autotune.SWEEP1, 1, None
        choice = getattr(self, "step_launch", None) or autotune.lookup(self._tune_key() + "/multi")
        if choice == 0:
            choice = autotune.SWEEP2 if min(self.nx, self.ny) >= 2 else autotune.SWEEP1
        m = autotune.steps_per_launch(choice)
        # the launchers are kept by runners the solver owns: a weak proxy keeps them
        # from forming a reference cycle with it
        me = weakref.proxy(self)
        if choice > autotune.TILE_BASE:
            nt = autotune.tile_threads(choice)
            return choice, m, (lambda s, d: me._stepn_into(s, d, m, threads=nt))
        if choice == autotune.SWEEP2:
            return choice, 2, (lambda s, d: me._step2_into(s, d))
        return choice, 1, None

    def _step_soa(self, f, ext_source=None, fpc=False, out=None):
        out = torch.empty_like(f) if out is None else out
        pc = torch.empty_like(f) if fpc else None
        nat.check(nat.lib().xb_darcy_step(self._args(f, out, ext_source, pc), nat.stream()), "xb_darcy_step")
        return out, pc

    def run_steps(self, f, n, chunk=64):
        """``n`` steps without I/O (the body of ``run`` when nothing happens between
        steps), replayed from a captured HIP graph of ``chunk`` steps."""
        fs = self._soa(f, 9)
        choice, m, multi = self._multi()
        key = ("steps", chunk, choice)
        runner = self._runners.get(key)
        if runner is None:
            me = weakref.proxy(self)
            runner = self._runners[key] = StepRunner(lambda s, d: me._step_soa(s, None, False, out=d), fs, chunk,
                                                     multi, m)
        return self._out(runner.run(fs, n))

    def run_steps_to(self, f, tol, check_every=10, max_steps=1000000, checks=10):
        """Steps from ``f`` until ``|f_k - f_{k-1}| < tol |f_1 - f_0|``, tested every
        ``check_every`` steps on the device (``k = 1 + j check_every``), the host
        reading the device-side latch once per ``checks`` tests.  Returns ``(f_k, k,
        ratio)``, ``k = -1`` if ``max_steps`` came first.  No source, no I/O."""
        fs = self._soa(f, 9)
        choice, m, multi = self._multi()
        key = ("steps_to", int(check_every), int(checks), choice)
        runner = self._runners.get(key)
        if runner is None:
            me = weakref.proxy(self)
            runner = self._runners[key] = ConvergedStepRunner(lambda s, d: me._step_soa(s, None, False, out=d), fs,
                                                              check_every, checks, multi, m)
        out, k, ratio = runner.run(fs, tol, max_steps)
        return self._out(out), k, ratio

    def collision(self, f, external_source=None):
        fs = self._soa(f, 9)
        out, pc = torch.empty_like(fs), torch.empty_like(fs)
        a = self._args(fs, out, self._soa(external_source, 1), pc)
        a.bc = None
        nat.check(nat.lib().xb_darcy_step(a, nat.stream()), "xb_darcy_step(collision)")
        return self._out(pc)

    def step(self, f_poststreaming, timestep, return_fpost=False, external_source=None):
        fs = self._soa(f_poststreaming, 9)
        out, pc = self._step_soa(fs, self._soa(external_source, 1), return_fpost)
        return self._out(out), self._out(pc)

    # ---- base.py:528-579 (zero-velocity equilibrium rho * w)
    def assign_fields_sharded(self):
        rho0, u0 = self.initialize_macroscopic_fields()
        w = torch.as_tensor(np.asarray(self.w, dtype=np.float64), dtype=self._tdt, device=self.device)
        if rho0 is None or u0 is None:
            return self._out(w.view(9, 1, 1).expand(9, self.nx, self.ny).contiguous())
        return self.initialize_populations(rho0, u0)

    def initialize_populations(self, rho0, u0):
        """``equilibrium(rho, u)`` (base.py:685-724); u is zero on this path."""
        rho = self._soa(rho0, 1)
        u = self._soa(u0, 2)
        if bool(torch.any(u != 0)):
            raise NotImplementedError("non-zero initial velocity (advection) is not supported on this path")
        f = self._new(9)
        w_host = np.asarray(self.w, dtype=np.float64)
        for i in range(9):
            nat.check(nat.lib().xb_lincomb(rho.numel(), nat.dtype_code(self._tdt), float(w_host[i]), nat.ptr(rho),
                                           0.0, None, nat.ptr(f[i]), nat.stream()), "xb_lincomb")
        return self._out(f)

    # ---- darcy.py:173-252 (t_max steps, from timestep 1)
    def run(self, t_max):
        f = self.assign_fields_sharded()
        if self.ioRate > 0:
            rho = io_field(self.update_macroscopic(f), self.downsamplingFactor)
            self.handle_io_timestep(0, f, None, rho, None)
        start = time.time()
        if self.ioRate == 0 and self.printInfoRate == 0 and not self.returnFpost and t_max >= 2:
            # nothing happens between steps: step 1 eagerly (the MLUPS clock starts after
            # it), the rest from captured graphs
            f, _ = self.step(f, 1)
            torch.cuda.synchronize()
            start = time.time()
            f = self.run_steps(f, t_max - 1)
            return self._finish_mlups(f, start, t_max)
        for timestep in range(1, t_max + 1):
            io_flag = self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max)
            print_iter_flag = self.printInfoRate > 0 and timestep % self.printInfoRate == 0
            if io_flag:
                rho_prev = io_field(self.update_macroscopic(f), self.downsamplingFactor)
            f, fstar = self.step(f, timestep, return_fpost=self.returnFpost)
            if print_iter_flag:
                print(f"Timestep {timestep} of {t_max} completed")
            if io_flag:
                rho = io_field(self.update_macroscopic(f), self.downsamplingFactor)
                self.handle_io_timestep(timestep, f, fstar, rho, rho_prev)
            if self.computeMLUPS and timestep == 1:
                torch.cuda.synchronize()
                start = time.time()
        return self._finish_mlups(f, start, t_max)

    def _finish_mlups(self, f, start, t_max):
        if self.computeMLUPS:
            torch.cuda.synchronize()
            end = time.time()
            self.mlups = self.nx * self.ny * t_max / (end - start) / 1e6
            print(f"Domain: {self.nx} x {self.ny}")
            print(f"Number of voxels: {self.nx * self.ny}")
            print(f"MLUPS: {self.mlups}")
        return f

    def handle_io_timestep(self, timestep, f, fstar, rho, rho_prev):
        self.output_data(timestep=timestep, rho=rho, rho_prev=rho_prev, f_poststreaming=f, f_postcollision=fstar)

    def get_source(self):
        pass


# ---------------------------------------------------------------------------
# ADR boundary conditions (darcy.py:298-466)

class ADR_AntiBounceBack(BoundaryCondition):
    """Halfway anti-bounce-back (Dirichlet ``rho``), post-streaming:
    ``f_missing = -f_known + 2 w rho`` (darcy.py:298-392)."""

    def __init__(self, indices, gridInfo, precision_policy, rho=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ADR_AntiBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.rho = rho

    def configure(self, boundaryMask):
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.rho is not None:
            if isinstance(self.rho, float) or np.asarray(self.rho).size == 1:
                # broadcast a scalar density (the reference only does so when the cell count changed)
                self.rho = np.full((nbd_modified, 1), float(np.asarray(self.rho).ravel()[0]))
            elif nbd_orig != nbd_modified:
                print("WARNING: Assuming a correct macroscopic (density) vector to impose at the BC cells!")
                self.rho = np.asarray(self.rho)[array_indices]

    def apply_rules(self, state):
        state.set_from_fin(self.imissing, self.iknown, -1)
        if self.rho is not None:
            w = np.asarray(self.lattice.w, dtype=np.float64)
            tmp = 2 * w * np.asarray(self.rho, dtype=np.float64).reshape(-1, 1)
            state.add(self.imissing, np.take_along_axis(tmp, np.asarray(self.iknown, dtype=np.int64), 1))


class ADR_Neumann_BounceBack(BounceBackHalfway):
    """Halfway bounce-back (no flux), post-streaming (darcy.py:394-415)."""

    def __init__(self, indices, gridInfo, precision_policy):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ADR_Neumann_BounceBack"


class ADR_EquilibriumBC(BoundaryCondition):
    """Cells set to ``rho w`` after streaming (darcy.py:417-466)."""

    def __init__(self, indices, gridInfo, precision_policy, rho, vel=None):
        super().__init__(indices, gridInfo, precision_policy)
        if vel is not None:
            raise NotImplementedError("ADR_EquilibriumBC with a velocity is not on this path")
        w = np.asarray(self.lattice.w, dtype=np.float64)
        rho = np.asarray(rho.detach().cpu() if isinstance(rho, torch.Tensor) else rho, dtype=np.float64)
        self.out = rho.reshape(-1, 1) * w if rho.ndim > 0 and rho.size > 1 else float(rho.ravel()[0]) * w
        self.name = "ADR_EquilibriumBC"
        self.implementationStep = "PostStreaming"

    def apply_rules(self, state):
        state.set_const(self.out)
