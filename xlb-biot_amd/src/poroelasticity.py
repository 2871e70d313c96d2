"""Biot poroelasticity: the D2Q8 elasticity solver coupled to the D2Q9 Darcy solver.

Drop-in for the reference ``src/poroelasticity.py`` (``LBMPoroelasticity``
:23-456): same constructor ``(LBMElasticityCls, LBMDarcyCls, **kwargs)``, keyword
arguments, validation, hooks (``get_elastic_force``, ``get_darcy_source``,
``initialize_macroscopic_fields``, ``output_data``) and ``step``/``run`` return
values.  The coupled step is a fixed sequence of GPU kernels over SoA buffers
(elementwise combinations, lattice gradients with the Darcy walls, the
elasticity steps or V-cycles with the implicit coupling force, the Darcy step).
"""
import time

import numpy as np
import torch

from . import _native as nat
from .base import as_soa, graph_capture, io_field, view


class BiotNonlinearForce:
    """``f_el -> darcy.gradient(elasticity.update_divergence(f_el), scale)``
    (poroelasticity.py:324), with a fused SoA form used inside the elasticity steps."""

    def __init__(self, elasticity, darcy, scale, fused=True):
        self.elasticity = elasticity
        self.darcy = darcy
        self.scale = scale
        # formed inside the elasticity sweeps from a carried divergence field
        self.fused = fused

    def __call__(self, f_el):
        return self.darcy.gradient(self.elasticity.update_divergence(f_el), self.scale)

    def soa_add(self, f_soa, base_soa):
        """``base + gradient(div(f), scale)`` on SoA buffers (``base`` may be None)."""
        div = self.elasticity._divergence_soa(f_soa)
        return self.darcy._gradient_soa(div, self.scale, base=base_soa, a=1.0, b=1.0)


class LBMPoroelasticity(object):
    def __init__(self, LBMElasticity, LBMDarcy, **kwargs):
        kwargs = dict(kwargs)
        kwargs_elasticity = kwargs.copy()
        kwargs_elasticity["lattice"] = kwargs.pop("lattice_elasticity")
        kwargs_elasticity["MG_levels"] = kwargs.pop("elasticity_MG_levels", 1)
        self.elasticity = LBMElasticity(**kwargs_elasticity)
        kwargs["lattice"] = kwargs.pop("lattice_darcy")
        self.darcy = LBMDarcy(**kwargs)

        self.elasticitySteps = kwargs.get("elasticity_steps", 1)
        self.alpha = kwargs.get("alpha")
        self.epsilon = kwargs.get("epsilon", 1 / self.elasticity.nx)
        self.couplingMethod = kwargs.get("coupling_method", "centered")
        self.alternative_gradient = kwargs.get("alternative_gradient", False)
        self.downsamplingFactor = kwargs.get("downsampling_factor", 1)
        self.printInfoRate = kwargs.get("print_info_rate", 100)
        self.ioRate = kwargs.get("io_rate", 0)
        self.computeMLUPS = kwargs.get("compute_MLUPS", False)
        if self.computeMLUPS:
            self.ioRate = 0
            self.printInfoRate = 0
        if (self.elasticity.nx, self.elasticity.ny) != (self.darcy.nx, self.darcy.ny):
            raise ValueError("elasticity and Darcy grids must coincide")
        if self.elasticity.precisionPolicy.torch_compute != self.darcy.precisionPolicy.torch_compute:
            raise ValueError("elasticity and Darcy must use the same precision")
        self.useDynamicForce = self.get_elastic_force(0) is not None
        self.useDynamicSource = self.get_darcy_source(0) is not None
        # extension: residual-driven inner loop (None keeps the reference's fixed count)
        self.elasticityTol = kwargs.get("elasticity_tol", None)
        self.elasticityMaxSteps = kwargs.get("elasticity_max_steps", 10000)
        # stop the residual-driven loop when the residual has not improved for this many
        # cycles (the reference V-cycle has no exact coarse solve and can stall or grow)
        self.elasticityPatience = kwargs.get("elasticity_patience", None)
        # "cycle": tolerance relative to the first residual of each inner solve;
        # "rhs": relative to the residual of the zero state under the step's force,
        # |S(0)| (a warm-started solve then stops above the round-off floor)
        self.elasticityTolScale = kwargs.get("elasticity_tol_scale", "cycle")
        if self.elasticityTolScale not in ("cycle", "rhs"):
            raise ValueError("elasticity_tol_scale must be 'cycle' or 'rhs'")
        self.inner_residuals = []
        self.useGraphs = kwargs.get("use_graphs", True)
        self.fusedCoupling = kwargs.get("fused_coupling", True)
        self._inner = None
        self.inner_counts = []
        self.mlups = None
        self.last_residual = None

    # ---- validated attributes (poroelasticity.py:72-176)
    @property
    def elasticitySteps(self):
        return self._elasticitySteps

    @elasticitySteps.setter
    def elasticitySteps(self, value):
        if not isinstance(value, int) or value < 1:
            raise TypeError("elasticitySteps must be a positive integer.")
        self._elasticitySteps = value

    @property
    def alpha(self):
        return self._alpha

    @alpha.setter
    def alpha(self, value):
        if not isinstance(value, float) or value < 0 or value > 1:
            raise TypeError("alpha must be a float between 0 and 1.")
        self._alpha = value

    @property
    def epsilon(self):
        return self._epsilon

    @epsilon.setter
    def epsilon(self, value):
        if not isinstance(value, float) or value <= 0:
            raise TypeError("epsilon must be a positive float.")
        self._epsilon = value

    @property
    def alternative_gradient(self):
        return self._alternative_gradient

    @alternative_gradient.setter
    def alternative_gradient(self, value):
        if not isinstance(value, bool):
            raise TypeError("alternative_gradient must be a boolean.")
        self._alternative_gradient = value

    @property
    def couplingMethod(self):
        return self._couplingMethod

    @couplingMethod.setter
    def couplingMethod(self, value):
        if value == "explicit":
            self.couplingRatio = 0.0
        elif value == "implicit":
            self.couplingRatio = 1.0
        elif value == "centered":
            self.couplingRatio = 0.5
        else:
            raise TypeError("couplingMethod must be 'explicit', 'implicit', 'centered', or 'alt_implicit', "
                            "'alt_centered'.")
        self._couplingMethod = value

    @property
    def couplingRatio(self):
        return self._couplingRatio

    @couplingRatio.setter
    def couplingRatio(self, value):
        if not isinstance(value, float) or value < 0.0 or value > 1.0:
            raise TypeError("couplingRatio must be a float beween 0 and 1")
        self._couplingRatio = value

    def _typed(name, typ, msg):
        def getter(self):
            return getattr(self, "_" + name)

        def setter(self, value):
            if not isinstance(value, typ):
                raise TypeError(msg)
            setattr(self, "_" + name, value)
        return property(getter, setter)

    downsamplingFactor = _typed("downsamplingFactor", int, "downsamplingFactor must be an integer")
    printInfoRate = _typed("printInfoRate", int, "printInfoRate must be an integer")
    ioRate = _typed("ioRate", int, "ioRate must be an integer")
    computeMLUPS = _typed("computeMLUPS", bool, "computeMLUPS must be a boolean")

    def show_simulation_parameters(self):
        for n in ["alpha", "couplingMethod", "elasticitySteps", "alternative_gradient", "epsilon",
                  "downsamplingFactor", "printInfoRate", "ioRate", "computeMLUPS"]:
            print(f"{n:>30} | {getattr(self, n, 'Attribute not set')}")

    # ---- helpers on SoA buffers
    @property
    def _tdt(self):
        return self.elasticity.precisionPolicy.torch_compute

    def _soa(self, x, c):
        if x is None:
            return None
        return as_soa(x, self.darcy.nx, self.darcy.ny, c, self._tdt, self.darcy.device)

    def _lincomb(self, a, x, b=0.0, y=None, out=None):
        out = torch.empty_like(x) if out is None else out
        nat.check(nat.lib().xb_lincomb(x.numel(), nat.dtype_code(x.dtype), float(a), nat.ptr(x), float(b),
                                       nat.ptr(y), nat.ptr(out), nat.stream()), "xb_lincomb")
        return out

    # ---- initialisation (poroelasticity.py:229-252)
    def initial_state(self):
        """``(f_elasticity, f_darcy, dt_div_eta)`` before the first step, as SoA buffers."""
        el, da = self.elasticity, self.darcy
        a, eps = self.alpha, self.epsilon
        dde = self.initialize_macroscopic_fields()
        if dde is None:
            dde = torch.zeros((1, da.nx, da.ny), dtype=self._tdt, device=da.device)
        else:
            dde = self._soa(dde, 1)
        fd = da._soa(da.assign_fields_sharded(), 9)
        src = self._lincomb(-a / eps, dde)
        if self.useDynamicSource:
            src = self._lincomb(1.0, src, 1.0, self._soa(self.get_darcy_source(0), 1))
        w = np.asarray(da.w, dtype=np.float64)
        fd_new = torch.empty_like(fd)
        for i in range(9):
            self._lincomb(1.0, fd[i], -0.5 * float(w[i]), src[0], out=fd_new[i])
        fd = fd_new
        pressure, _ = da._macro_soa(fd, src, True, False)
        fe = el._soa(el.assign_fields_sharded(), 8)
        if self.alternative_gradient:
            force = da._gradient_soa(pressure, -a * eps)
        else:
            _, pg = da._macro_soa(fd, None, False, True)
            pg = da._gradient_soa(src, 0.5, base=pg, a=1.0, b=1.0)
            force = self._lincomb(-a * eps, pg)
        if self.useDynamicForce:
            force = self._lincomb(1.0, force, 1.0, self._soa(self.get_elastic_force(0), 2))
        eta, sigma = el.initialize_macroscopic_fields()
        if eta is not None and sigma is not None:
            fe = el._soa(el.equilibrium(eta, sigma, external_force=view(force)), 8)
        return fe, fd, dde

    # ---- coupled step (poroelasticity.py:303-339)
    def _step_soa(self, fe, fd, dde, timestep):
        ctx = self._pre_inner(fe, fd, dde, timestep)
        fe = self._inner_solve(fe, ctx["force"], ctx["nlf"])
        return self._post_inner(ctx, fe)

    def coupling_force(self):
        """The solution-dependent part of the elastic force, ``-gradient(div f, s)``
        with ``s = theta alpha^2 / 2`` (poroelasticity.py:324)."""
        s_nl = 0.5 * self.couplingRatio * self.alpha ** 2
        return BiotNonlinearForce(self.elasticity, self.darcy, s_nl, fused=self.fusedCoupling)

    def _pre_inner(self, fe, fd, dde, timestep):
        """Everything of the coupled step before the elasticity iterations
        (poroelasticity.py:303-324): the static part of the elastic force."""
        el, da = self.elasticity, self.darcy
        th, a, eps = self.couplingRatio, self.alpha, self.epsilon
        source_base = self._lincomb((1.0 - th) * -a / eps, dde)
        if self.useDynamicSource:
            source_base = self._lincomb(1.0, source_base, 1.0, self._soa(self.get_darcy_source(timestep), 1))
        if self.alternative_gradient:
            p_base, _ = da._macro_soa(fd, source_base, True, False)
            pgb = da._gradient_soa(p_base, 1.0)
        else:
            _, g = da._macro_soa(fd, None, False, True)
            pgb = da._gradient_soa(source_base, 0.5, base=g, a=1.0, b=1.0)
        nlf = self.coupling_force()
        force = da._gradient_soa(el._divergence_soa(fe), nlf.scale, base=pgb, a=-a * eps, b=-1.0)
        if self.useDynamicForce:
            force = self._lincomb(1.0, force, 1.0, self._soa(self.get_elastic_force(timestep), 2))
        return dict(fprev=fe, fd=fd, source_base=source_base, pgb=pgb, force=force, nlf=nlf)

    def _post_inner(self, ctx, fe):
        """The rest of the coupled step (poroelasticity.py:326-339) from the new
        elastic populations: stresses, divergence rate, pressure, Darcy step."""
        el, da = self.elasticity, self.darcy
        th, a, eps = self.couplingRatio, self.alpha, self.epsilon
        fprev, fd, source_base, pgb = ctx["fprev"], ctx["fd"], ctx["source_base"], ctx["pgb"]
        ext = ctx["nlf"].soa_add(fe, ctx["force"])
        eta = el._new(2)
        sigma = el._new(3)
        nat.check(nat.lib().xb_elast_macro(el.nx, el.ny, nat.dtype_code(self._tdt), el.lamda, el.mu, nat.ptr(fe),
                                           nat.ptr(el._force_soa), nat.ptr(ext), nat.ptr(eta), nat.ptr(sigma),
                                           *el._fields(), nat.stream()), "xb_elast_macro")
        dde = el._divergence_soa(fe, fprev)
        source = self._lincomb(1.0, source_base, -(th * a / eps), dde)
        pressure, _ = da._macro_soa(fd, source, True, False)
        if self.alternative_gradient:
            pressure_grad = da._gradient_soa(pressure, 1.0)
        else:
            pressure_grad = da._gradient_soa(dde, 0.5 * th * a / eps, base=pgb, a=1.0, b=-1.0)
        fd, _ = da._step_soa(fd, source)
        return fe, fd, eta, sigma, dde, pressure, pressure_grad

    def prepare_inner(self):
        """Capture the graphs of the inner elasticity iterations now (they are
        otherwise captured in the first step)."""
        if not self.useGraphs:
            return
        el = self.elasticity
        nlf = self.coupling_force()
        force = torch.zeros((2, el.nx, el.ny), dtype=self._tdt, device=el.device)
        self._inner = el.vcycle_graph(force, nlf) if el.levels > 1 else _PlainInnerGraph(el, force, nlf)
        self._inner_key = (nlf.scale, nlf.fused)

    def probe_levels(self, cycles=60):
        """Residual norms of the first ``cycles`` V-cycles of the first physical step
        and the seconds per cycle (for :func:`autotune.tune_levels`)."""
        from .autotune import probe_vcycle
        fe, fd, dde = self.initial_state()
        ctx = self._pre_inner(fe, fd, dde, 0)
        vg = self.elasticity.vcycle_graph(ctx["force"], ctx["nlf"])
        return probe_vcycle(vg, fe, cycles)

    def _inner_solve(self, fe, force, nlf):
        """The elasticity iterations of one physical step (poroelasticity.py:325-326):
        ``elasticity_steps`` steps / V-cycles, or, with ``elasticity_tol``, V-cycles
        until the residual norm falls below ``tol`` times that of the first cycle."""
        el = self.elasticity
        tol = self.elasticityTol
        if tol is not None and el.levels < 2:
            raise ValueError("elasticity_tol needs a multigrid elasticity solver (elasticity_MG_levels >= 2)")
        if not self.useGraphs:
            n = self.elasticitySteps if tol is None else self.elasticityMaxSteps
            r0 = None
            for k in range(n):
                fe, aux = el._step_soa(fe, force, nlf)
                if aux is not None:
                    self.last_residual = aux
                    if tol is not None:
                        r = aux.item()
                        if not np.isfinite(r):
                            raise FloatingPointError(f"V-cycle residual is not finite after {k + 1} cycles")
                        r0 = r if r0 is None else r0
                        if r <= tol * r0:
                            break
            self.inner_counts.append(k + 1)
            return fe
        if self._inner is None or self._inner_key != (nlf.scale, nlf.fused):
            self._inner = (el.vcycle_graph(force, nlf) if el.levels > 1 else _PlainInnerGraph(el, force, nlf))
            self._inner_key = (nlf.scale, nlf.fused)
        inner = self._inner
        inner.set_force(force)
        inner.load(fe)
        if tol is None:
            inner.run(self.elasticitySteps)
            count = self.elasticitySteps
        else:
            r_ref = None
            if self.elasticityTolScale == "rhs":
                r_ref = self.elasticity.rhs_residual_norm(force, nlf).item()
            count, rel, _ = inner.run_to(tol, self.elasticityMaxSteps, self.elasticityPatience, r_ref)
            self.inner_residuals.append(rel)
        self.inner_counts.append(count)
        self.last_residual = getattr(inner, "norm", None)
        return inner.state.clone()

    def step(self, f_elasticity, f_darcy, dt_div_eta, timestep):
        fe = self.elasticity._soa(f_elasticity, 8)
        fd = self.darcy._soa(f_darcy, 9)
        dde = self._soa(dt_div_eta, 1)
        out = self._step_soa(fe, fd, dde, timestep)
        # populations in the storage precision, macroscopic fields in the compute one
        return (self.elasticity._out(out[0]), self.darcy._out(out[1])) + tuple(view(t) for t in out[2:])

    # ---- driver (poroelasticity.py:208-301): t_max + 1 physical steps
    def run(self, t_max):
        fe, fd, dde = self.initial_state()
        start = time.time()
        for timestep in range(0, t_max + 1):
            io_flag = self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max)
            print_iter_flag = self.printInfoRate > 0 and timestep % self.printInfoRate == 0
            fe, fd, eta, sigma, dde, pressure, pressure_grad = self._step_soa(fe, fd, dde, timestep)
            if print_iter_flag:
                print(f"Timestep {timestep} of {t_max} completed")
            if io_flag:
                g = lambda x: io_field(view(x), self.downsamplingFactor)
                self.handle_io_timestep(timestep, self.elasticity._out(fe), self.darcy._out(fd), g(eta), g(sigma),
                                        g(pressure), g(pressure_grad), g(dde))
            if self.computeMLUPS and timestep == 1:
                torch.cuda.synchronize()
                start = time.time()
        if self.computeMLUPS:
            torch.cuda.synchronize()
            end = time.time()
            nx, ny = self.elasticity.nx, self.elasticity.ny
            self.mlups = nx * ny * t_max / (end - start) / 1e6
            print(f"Domain: {nx} x {ny}")
            print(f"Number of voxels: {nx * ny}")
            print(f"MLUPS: {self.mlups}")
        return self.elasticity._out(fe), self.darcy._out(fd)

    def handle_io_timestep(self, timestep, f_elasticity, f_darcy, eta, sigma, pressure, pressure_grad, dt_div_eta):
        self.output_data(timestep=timestep, eta=eta, sigma=sigma, dt_div_eta=dt_div_eta, pressure=pressure,
                         pressure_grad=pressure_grad, f_poststreaming_elasticity=f_elasticity,
                         f_poststreaming_darcy=f_darcy)

    # ---- hooks (poroelasticity.py:341-456)
    def output_data(self, **kwargs):
        pass

    def get_elastic_force(self, timestep):
        pass

    def get_darcy_source(self, timestep):
        pass

    def initialize_macroscopic_fields(self):
        print("WARNING: Default initial conditions assumed: dt_div_eta = 0")
        print("         To set explicit initial dt_div_eta, use self.initialize_macroscopic_fields.")
        return None


class _PlainInnerGraph:
    """Single-grid inner iterations (coupling force + step) replayed from a captured
    HIP graph of ``chunk`` steps; the remainder runs eagerly.  With the fused coupling
    force the divergence of the state is carried in ``da`` / ``db``."""

    def __init__(self, el, force, nlf, chunk=16):
        self.el, self.nlf, self.chunk = el, nlf, chunk
        self.fused = el._fusable(nlf)
        self.ext = force.clone()
        self.a = torch.zeros((8, el.nx, el.ny), dtype=el._tdt, device=el.device)
        self.b = torch.empty_like(self.a)
        self.da = torch.zeros((1, el.nx, el.ny), dtype=el._tdt, device=el.device)
        self.db = torch.empty_like(self.da)
        self._one(self.a, self.da, self.b, self.db)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with graph_capture(self.graph):
            for _ in range(chunk // 2):
                self._one(self.a, self.da, self.b, self.db)
                self._one(self.b, self.db, self.a, self.da)

    def _one(self, s, ds, d, dd):
        if self.fused:
            self.el._step_nl(s, ds, 0, self.ext, self.nlf, out=d, out_div=dd)
        else:
            self.el._step_into(s, d, self.nlf.soa_add(s, self.ext))

    @property
    def state(self):
        return self.a

    def load(self, f):
        self.a.copy_(f)
        if self.fused:
            el = self.el
            nat.check(nat.lib().xb_elast_divergence(el.nx, el.ny, nat.dtype_code(f.dtype), el.lamda, el.mu,
                                                    nat.ptr(self.a), None, nat.ptr(self.da), *el._fields(), nat.stream()),
                      "xb_elast_divergence")

    def set_force(self, force):
        self.ext.copy_(force)

    def run(self, n):
        done = 0
        while n - done >= self.chunk:
            self.graph.replay()
            done += self.chunk
        while done < n:
            self._one(self.a, self.da, self.b, self.db)
            self.a.copy_(self.b)
            self.da.copy_(self.db)
            done += 1
