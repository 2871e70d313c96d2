"""D2Q8 lattice-Boltzmann solver for static linear elasticity, with the multigrid V-cycle.

Drop-in for the reference ``src/elasticity.py`` (``LBMElasticity`` :30-1501 and the
elastic boundary conditions :1503-1750): same class names, keyword arguments,
validation errors, hooks (``set_boundary_conditions``, ``get_force``,
``initialize_macroscopic_fields``, ``output_data``) and ``step``/``run``
signatures and return values.  Every step runs on the GPU through the C ABI of
``include/xlb_biot.h``; arrays are torch tensors shaped like the reference's.
"""
import ctypes
import time
import weakref

import numpy as np
import torch

from . import _native as nat
from . import autotune
from .base import (ConvergedStepRunner, NormWorkspace, StepRunner, PrecisionPolicy, as_soa, default_device, graph_capture, io_field, no_capture,
                   torch_dtype, view)
from .boundary_conditions import BoundaryCondition, build_rule_table, configure_bcs, create_grid_mask


class LBMElasticity(object):
    """Moment-based D2Q8 LBM for static linear elasticity (reference elasticity.py:30-148)."""

    def __init__(self, **kwargs):
        self.lamda = kwargs.get("lambda")
        self.mu = kwargs.get("mu")
        self.nx = kwargs.get("nx")
        self.ny = kwargs.get("ny")
        self.nz = kwargs.get("nz")
        self.mg_BCs = kwargs.get("MG_BCs", {})
        self.levels = kwargs.get("MG_levels", 1)
        self.gamma = kwargs.get("MG_relaxation", 0.75)
        self.k_smooth = kwargs.get("MG_n_smoothing", 2)
        self.precision = kwargs.get("precision")
        self.precisionPolicy = PrecisionPolicy(self.precision)
        if self.precisionPolicy.mixed:
            # narrower storage is rounded by the single sweeps: no multi-step smoothing
            # tiles, no single-workgroup coarse tail
            self.use_tail = False
            self.smooth_tile_steps = 1
        self.lattice = kwargs.get("lattice")
        self.downsamplingFactor = kwargs.get("downsampling_factor", 1)
        self.printInfoRate = kwargs.get("print_info_rate", 100)
        self.ioRate = kwargs.get("io_rate", 0)
        self.returnFpost = kwargs.get("return_fpost", False)
        self.computeMLUPS = kwargs.get("compute_MLUPS", False)
        self.nDevices = 1
        self.backend = "rocm"
        self.device = default_device()
        self.sharding = None
        if self.computeMLUPS:
            self.ioRate = 0
            self.printInfoRate = 0
        self.c = self.lattice.c
        self.q = self.lattice.q
        self.w = self.lattice.w
        self.dim = self.lattice.d
        self._create_transfer_matrices()
        self.gridInfo = {"nx": self.nx, "ny": self.ny, "nz": self.nz, "dim": self.lattice.d, "lattice": self.lattice}
        self.boundingBoxIndices = self.bounding_box_indices()
        self._create_boundary_data()
        force = self.get_force()
        self._force_soa = None if force is None else as_soa(force, self.nx, self.ny, 2, self._tdt, self.device)
        self.force = None if self._force_soa is None else view(self._force_soa)
        # extension: per-cell Lame parameters (nx, ny[, 1]) of the finest grid; they
        # replace lambda / mu in its collision, stresses, divergence and equilibrium.
        # Coarse multigrid levels keep the scalar lambda / mu (their operator only
        # preconditions the fine-grid fixed point).
        self._lam_field = self._mu_field = None
        lf, mf = kwargs.get("lambda_field"), kwargs.get("mu_field")
        if (lf is None) != (mf is None):
            raise ValueError("lambda_field and mu_field go together")
        if lf is not None:
            self.set_lame_fields(lf, mf)
        self._norm_ws = {}
        self._runners = {}
        self.mlups = None
        self.last_residual = None

    def set_lame_fields(self, lam_field, mu_field):
        """Per-cell Lame parameters ``(nx, ny[, 1])`` of the finest grid (arrays or
        device tensors, copied; a second call overwrites the first fields in place)."""
        lf = as_soa(lam_field, self.nx, self.ny, 1, self._tdt, self.device)
        mf = as_soa(mu_field, self.nx, self.ny, 1, self._tdt, self.device)
        if self._lam_field is not None:
            # in place: captured step and V-cycle graphs hold these addresses
            self._lam_field.copy_(lf)
            self._mu_field.copy_(mf)
            return
        self._lam_field, self._mu_field = lf.clone(), mf.clone()
        self._runners = {}

    def _fields(self):
        return nat.ptr(self._lam_field), nat.ptr(self._mu_field)

    def _drop_plans(self):
        """Forget captured step loops, coarse-level tails and adjoint sessions: they
        hold lambda, mu, gamma and k_smooth by value (called by those setters)."""
        for name in ("_runners", "_tails", "_adjoint_sessions"):
            if name in self.__dict__:
                self.__dict__[name] = {}

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

    # ---- validated attributes (elasticity.py:150-354)
    @property
    def lattice(self):
        return self._lattice

    @lattice.setter
    def lattice(self, value):
        if value is None:
            raise ValueError("Lattice type must be provided.")
        if self.nz == 0 and value.name not in ["D2Q8"]:
            raise ValueError("For 2D simulations, lattice type must be LatticeD2Q8.")
        if self.nz != 0:
            raise ValueError("3D simulations are not supported.")
        self._lattice = value

    @property
    def lamda(self):
        return self._lamda

    @lamda.setter
    def lamda(self, value):
        if value is None:
            raise ValueError("lambda must be provided")
        if not isinstance(value, float):
            raise TypeError("lambda must be a float")
        self._lamda = value
        self._drop_plans()

    @property
    def mu(self):
        return self._mu

    @mu.setter
    def mu(self, value):
        if value is None:
            raise ValueError("mu must be provided")
        if not isinstance(value, float):
            raise TypeError("mu must be a float")
        self._mu = value
        self._drop_plans()

    @property
    def levels(self):
        return self._levels

    @levels.setter
    def levels(self, value):
        if value is None:
            raise ValueError("levels must be provided")
        if not isinstance(value, int) or value <= 0:
            raise TypeError("levels must be a positive int")
        den = 2 ** (value - 1)
        if (self.nx - int(self.mg_BCs["left"]) - int(self.mg_BCs["right"])) % den != 0 \
                or (self.ny - int(self.mg_BCs["bottom"]) - int(self.mg_BCs["top"])) % den != 0 \
                or (self.nz - int(self.mg_BCs["front"]) - int(self.mg_BCs["back"])) % den != 0:
            raise ValueError("nx,ny,nz (-BCs) must be divisible by 2**(levels-1)")
        self._levels = value

    @property
    def gamma(self):
        return self._gamma

    @gamma.setter
    def gamma(self, value):
        if value is None:
            raise ValueError("gamma must be provided")
        if not isinstance(value, float):
            raise TypeError("gamma must be a float")
        if value < 0.0 or value > 1.0:
            raise ValueError("gamma must be between 0 and 1")
        self._gamma = value
        self._drop_plans()

    @property
    def k_smooth(self):
        return self._k_smooth

    @k_smooth.setter
    def k_smooth(self, value):
        if value is None:
            raise ValueError("k_smooth must be provided")
        if not isinstance(value, int):
            raise TypeError("k_smooth must be an integer")
        if value < 1:
            raise ValueError("k_smooth must be positive")
        self._k_smooth = value
        self._drop_plans()

    @property
    def mg_BCs(self):
        return self._mg_BCs

    @mg_BCs.setter
    def mg_BCs(self, value):
        if not isinstance(value, dict):
            raise TypeError("mg_BCs must be a dictionary")
        for key, bc in value.items():
            if key not in ["left", "right", "bottom", "top", "front", "back"]:
                raise ValueError(f"Invalid key '{key}' in mg_BCs. Valid keys are 'left', 'right', 'bottom', 'top', "
                                 "'front', 'back'.")
            if not isinstance(bc, bool):
                raise TypeError("All values in mg_BCs must be boolean values")
        self._mg_BCs = {"left": False, "right": False, "bottom": False, "top": False, "front": False, "back": False}
        self._mg_BCs.update(value)

    def _int_attr(name):
        def getter(self):
            return getattr(self, "_" + name)

        def setter(self, value):
            if value is None:
                raise ValueError(f"{name} must be provided")
            if not isinstance(value, int):
                raise TypeError(f"{name} must be an integer")
            setattr(self, "_" + name, value)
        return property(getter, setter)

    nx = _int_attr("nx")
    ny = _int_attr("ny")
    nz = _int_attr("nz")

    def _typed(name, typ, msg):
        def getter(self):
            return getattr(self, "_" + name)

        def setter(self, value):
            if not isinstance(value, typ):
                raise TypeError(msg)
            setattr(self, "_" + name, value)
        return property(getter, setter)

    precision = _typed("precision", str, "precision must be a string")
    downsamplingFactor = _typed("downsamplingFactor", int, "downsamplingFactor must be an integer")
    printInfoRate = _typed("printInfoRate", int, "printInfoRate must be an integer")
    ioRate = _typed("ioRate", int, "ioRate must be an integer")
    returnFpost = _typed("returnFpost", bool, "returnFpost must be a boolean")
    computeMLUPS = _typed("computeMLUPS", bool, "computeMLUPS must be a boolean")
    nDevices = _typed("nDevices", int, "nDevices must be an integer")

    def show_simulation_parameters(self):
        names = ["lamda", "mu", "nx", "ny", "nz", "dim", "levels", "gamma", "k_smooth", "precision", "lattice",
                 "downsamplingFactor", "printInfoRate", "ioRate", "computeMLUPS", "backend", "nDevices"]
        print(f"**** Simulation Parameters for {self.__class__.__name__} ****")
        for n in names:
            print(f"{n:>30} | {getattr(self, n, 'Attribute not set')}")

    # ---- geometry (elasticity.py:501-604)
    def grid_size(self, level=0):
        if level == 0:
            return self.nx, self.ny
        nx_bc = int(self.mg_BCs["left"]) + int(self.mg_BCs["right"])
        ny_bc = int(self.mg_BCs["bottom"]) + int(self.mg_BCs["top"])
        return (self.nx - nx_bc) // 2 ** level + nx_bc, (self.ny - ny_bc) // 2 ** level + ny_bc

    def grid_info(self, level=0):
        nx, ny = self.grid_size(level)
        return {"nx": nx, "ny": ny, "nz": self.nz, "dim": self.lattice.d, "lattice": self.lattice}

    def bounding_box_indices(self, level=0):
        nx, ny = self.grid_size(level)
        return {"bottom": np.array([[i, 0] for i in range(nx)], dtype=int),
                "top": np.array([[i, ny - 1] for i in range(nx)], dtype=int),
                "left": np.array([[0, i] for i in range(ny)], dtype=int),
                "right": np.array([[nx - 1, i] for i in range(ny)], dtype=int)}

    def interior_box(self, level=0):
        """Cells outside the multigrid BC layers: ``(x0, x1, y0, y1)``."""
        nx, ny = self.grid_size(level)
        return (int(self.mg_BCs["left"]), nx - int(self.mg_BCs["right"]),
                int(self.mg_BCs["bottom"]), ny - int(self.mg_BCs["top"]))

    def set_precisions(self, precision):
        return PrecisionPolicy.TABLE.get(precision, (np.float32, np.float32))

    def create_grid_mask(self, solid_halo_voxels, level=0):
        nx, ny = self.grid_size(level)
        return create_grid_mask(nx, ny, self.lattice, solid_halo_voxels)

    def _create_boundary_data(self):
        # elasticity.py:396-436; then fold each level's BCs into a kernel rule table
        self.BCs = []
        self.BCs_MG = {level: [] for level in range(1, self.levels)}
        self.set_boundary_conditions()
        start = time.time()
        self._tables = {}
        for level in range(self.levels):
            bcs = self.BCs if level == 0 else self.BCs_MG[level]
            nx, ny = self.grid_size(level)
            configure_bcs(bcs, nx, ny, self.lattice)
            self._tables[level] = build_rule_table(bcs, nx, ny, self.q, self._tdt, self.device)
        self._bc_build_time = time.time() - start

    def distributed_array_init(self, shape, type, init_val=0, sharding=None):
        dt = torch_dtype(type)
        if isinstance(init_val, torch.Tensor):
            v = init_val.to(device=self.device, dtype=dt)
        else:
            v = torch.as_tensor(np.asarray(init_val), dtype=dt, device=self.device)
        return v.expand(shape).clone() if tuple(v.shape) != tuple(shape) else v.clone()

    # ---- hooks (elasticity.py:631-645, 1458-1501)
    def initialize_macroscopic_fields(self):
        print("WARNING: Default initial conditions assumed: displacement = 0, stress = 0")
        print("         To set explicit initial displacement and stress, use self.initialize_macroscopic_fields.")
        return None, None

    def output_data(self, **kwargs):
        pass

    def set_boundary_conditions(self):
        pass

    def get_force(self):
        pass

    # ---- constants (elasticity.py:805-848); host copies, the kernels form the same entries
    def _create_transfer_matrices(self):
        g = 1.0 / (12.0 * (self.lamda + self.mu) - 4.0)
        if abs(g) > 100:
            print("Warning: scaling value is too large. Reset last moment to default. Boundary conditions may fail.")
            g = 0.0
        self.f_to_m = np.array([[1, 0, -1, 0, 1, -1, -1, 1],
                                [0, 1, 0, -1, 1, 1, -1, -1],
                                [1, 1, 1, 1, 2, 2, 2, 2],
                                [1, -1, 1, -1, 0, 0, 0, 0],
                                [0, 0, 0, 0, 1, -1, 1, -1]], dtype=np.float64).T
        h, n = 0.25 + 0.5 * g, -0.25 * g
        self.m_to_f = np.array([[0.5, 0, h, 0.25, 0, -0.5, 0],
                                [0, 0.5, h, -0.25, 0, 0, -0.5],
                                [-0.5, 0, h, 0.25, 0, 0.5, 0],
                                [0, -0.5, h, -0.25, 0, 0, 0.5],
                                [0, 0, n, 0, 0.25, 0.25, 0.25],
                                [0, 0, n, 0, -0.25, -0.25, 0.25],
                                [0, 0, n, 0, 0.25, -0.25, -0.25],
                                [0, 0, n, 0, -0.25, 0.25, -0.25]], dtype=np.float64).T

    # ---- conversions
    def _soa(self, x, c, level=0):
        if x is None:
            return None
        nx, ny = self.grid_size(level)
        return as_soa(x, nx, ny, c, self._tdt, self.device)

    def _new(self, c, level=0):
        nx, ny = self.grid_size(level)
        return torch.empty((c, nx, ny), dtype=self._tdt, device=self.device)

    # ---- kernels
    def _tune_key(self, level, mode, has_ext, has_res, two=False):
        nx, ny = self.grid_size(level)
        return (f"elast/{self.precision}/{nx}x{ny}/mode{mode}/bc{int(self._tables.get(level) is not None)}"
                f"/ext{int(has_ext)}/res{int(has_res)}" + ("/x2" if two else ""))

    def _launch(self, f_in, f_out, level, mode, ext=None, resid=None, fpc=None, coarse=None, partials=None,
                use_force=None):
        # f_in None: the zero state (tile smoothing only; the shape comes from f_out)
        ref = f_in if f_in is not None else f_out
        nx, ny = ref.shape[1], ref.shape[2]
        x0, x1, y0, y1 = self.interior_box(level)
        table = self._tables.get(level)
        if use_force is None:
            use_force = level == 0
        fa = self._force_soa if use_force else None
        a = nat.ElastArgs(nx, ny, self._code(ref), mode, self.lamda, self.mu, self.gamma,
                          nat.ptr(f_in), nat.ptr(f_out), nat.ptr(fpc), nat.ptr(fa), nat.ptr(ext), nat.ptr(resid),
                          x0, x1, y0, y1, table.ref() if table is not None else None,
                          nat.ptr(coarse), nat.ptr(partials),
                          autotune.lookup(self._tune_key(level, mode, ext is not None, resid is not None)), 0)
        if level == 0 and self._lam_field is not None:
            a.lam_field, a.mu_field = self._fields()
        return a

    def autotune(self, levels=None, external_force=False):
        """Measure and cache the best sweep launch shape for the step / smoothing /
        residual kernels of every level of this solver (with ``external_force``, for
        the finest-level launches that carry an external force, as in the Biot step)."""
        lib = nat.lib()
        out = {}
        for level in (range(self.levels) if levels is None else levels):
            nx, ny = self.grid_size(level)
            f = torch.zeros((8, nx, ny), dtype=self._tdt, device=self.device)
            o = torch.empty_like(f)
            x0, x1, y0, y1 = self.interior_box(level)
            r = torch.zeros((8, x1 - x0, y1 - y0), dtype=self._tdt, device=self.device) if level > 0 else None
            modes = [nat.XB_MODE_STEP] if self.levels == 1 else [nat.XB_MODE_SMOOTH]
            if level < self.levels - 1:
                modes.append(nat.XB_MODE_RESIDUAL)
            ext = (torch.zeros((2, nx, ny), dtype=self._tdt, device=self.device)
                   if external_force and level == 0 else None)
            for mode in modes:
                coarse = torch.empty((8, (x1 - x0) // 2, (y1 - y0) // 2), dtype=self._tdt, device=self.device) \
                    if mode == nat.XB_MODE_RESIDUAL else None
                a = self._launch(f, o, level, mode, ext=ext, resid=r, coarse=coarse)
                a.rows = 2  # the most wavefronts, hence the most partial sums
                ws = NormWorkspace(lib.xb_elast_partials_count(a), self.device)
                a.norm_partials = nat.ptr(ws.partials)

                def run(rows, a=a):
                    a.rows = rows
                    nat.check(lib.xb_elast_step(a, nat.stream()), "xb_elast_step(tune)")
                key = self._tune_key(level, mode, ext is not None, r is not None)
                cands = [c for c in autotune.CANDIDATES if c >= 2] if mode == nat.XB_MODE_RESIDUAL \
                    else autotune.CANDIDATES
                out[key] = autotune.tune(key, run, cands)
                if mode == nat.XB_MODE_STEP and min(nx, ny) >= 2:
                    a2 = self._launch(f, o, level, mode, ext=ext)

                    def run2(rows, a2=a2):
                        a2.rows = rows
                        nat.check(lib.xb_elast_step2(a2, nat.stream()), "xb_elast_step2(tune)")
                    key2 = self._tune_key(level, mode, ext is not None, False, True)
                    out[key2] = autotune.tune(key2, run2)
        if self.levels == 1 and (levels is None or 0 in levels):
            out.update(self.tune_step_loops(external_force))
        return out

    def _run_step(self, f, ext=None, level=0, fpc=False):
        """One ``_step`` on SoA buffers: returns ``(f_new, f_postcollision | None)``."""
        out = torch.empty_like(f)
        pc = torch.empty_like(f) if fpc else None
        a = self._launch(f, out, level, nat.XB_MODE_STEP, ext=ext, fpc=pc)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step")
        return out, pc

    def _step_into(self, f, out, ext=None, level=0):
        a = self._launch(f, out, level, nat.XB_MODE_STEP, ext=ext)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step")
        return out

    def _step2_into(self, f, out, ext=None):
        """Two single-grid steps in one pass (``xb_elast_step2``)."""
        a = self._launch(f, out, 0, nat.XB_MODE_STEP, ext=ext)
        a.rows = autotune.lookup(self._tune_key(0, nat.XB_MODE_STEP, ext is not None, False, True))
        nat.check(nat.lib().xb_elast_step2(a, nat.stream()), "xb_elast_step2")
        return out

    def _stepn_into(self, f, out, m, ext=None, threads=0):
        """``m`` single-grid steps in one launch on LDS tiles (``xb_elast_stepn``) of
        ``threads`` threads per block (0: the library's choice)."""
        a = self._launch(f, out, 0, nat.XB_MODE_STEP, ext=ext)
        a.rows = int(threads)
        nat.check(nat.lib().xb_elast_stepn(a, int(m), nat.stream()), "xb_elast_stepn")
        return out

    def _multi_key(self, has_ext):
        return self._tune_key(0, nat.XB_MODE_STEP, has_ext, False) + "/multi"

    def tune_step_loops(self, external_force=False):
        """Measure and cache the best number of single-grid steps per launch for step
        loops: one or two per sweep, or 2 .. 8 on LDS tiles."""
        f = torch.zeros((8, self.nx, self.ny), dtype=self._tdt, device=self.device)
        o = torch.empty_like(f)
        ext = torch.zeros((2, self.nx, self.ny), dtype=self._tdt, device=self.device) if external_force else None
        launches = {autotune.SWEEP1: lambda: self._step_into(f, o, ext)}
        if min(self.nx, self.ny) >= 2:
            launches[autotune.SWEEP2] = lambda: self._step2_into(f, o, ext)
        for m in autotune.TILE_STEPS:
            for nt in autotune.TILE_THREADS:
                launches[autotune.tile_choice(m, nt)] = (lambda m=m, nt=nt: self._stepn_into(f, o, m, ext, nt))
        key = self._multi_key(external_force)
        return {key: autotune.tune_multi(key, launches)}

    def _multi(self, ext):
        if self.precisionPolicy.mixed:
            # narrower storage is rounded by the single-step sweep only
            return autotune.SWEEP1, 1, None
        choice = autotune.lookup(self._multi_key(ext is not None))
        if choice == 0:
            choice = autotune.SWEEP2 if min(self.nx, self.ny) >= 2 else autotune.SWEEP1
        m = autotune.steps_per_launch(choice)
        # the launchers are kept by runners the solver owns: a weak proxy keeps them
        # from forming a reference cycle with it
        me = weakref.proxy(self)
        if choice > autotune.TILE_BASE:
            nt = autotune.tile_threads(choice)
            return choice, m, (lambda s, d: me._stepn_into(s, d, m, ext, nt))
        if choice == autotune.SWEEP2:
            return choice, 2, (lambda s, d: me._step2_into(s, d, ext))
        return choice, 1, None

    def run_steps(self, f, n, chunk=64, external_force=None):
        """``n`` single-grid steps with a fixed external force and no I/O, replayed
        from a captured HIP graph of about ``chunk`` steps (the tuned number of steps
        per launch, see :meth:`tune_step_loops`)."""
        if self.levels != 1:
            raise ValueError("run_steps is the single-grid loop; multigrid uses step() / run()")
        fs = self._soa(f, 8)
        ext = self._soa(external_force, 2)
        choice, m, multi = self._multi(ext)
        key = ("steps", chunk, external_force is None, choice)
        runner = self._runners.get(key)
        if runner is None or external_force is not None:
            me = weakref.proxy(self)
            runner = StepRunner(lambda s, d: me._step_into(s, d, ext), fs, chunk, multi, m)
            if external_force is None:
                self._runners[key] = runner
        return self._out(runner.run(fs, n))

    def run_steps_to(self, f, tol, check_every=10, max_steps=1000000, checks=10):
        """Single-grid steps from ``f`` until ``|f_k - f_{k-1}| < tol |f_1 - f_0|``, tested
        every ``check_every`` steps on the device (``k = 1 + j check_every``; the test
        and the state at the first such ``k`` are latched on the device and the host
        reads the latch once per ``checks`` tests).  Returns ``(f_k, k, ratio)``, with
        ``k = -1`` if ``max_steps`` came first.  No external force, no I/O."""
        if self.levels != 1:
            raise ValueError("run_steps_to is the single-grid loop; multigrid uses step() / run()")
        fs = self._soa(f, 8)
        choice, m, multi = self._multi(None)
        key = ("steps_to", int(check_every), int(checks), choice)
        runner = self._runners.get(key)
        if runner is None:
            me = weakref.proxy(self)
            runner = ConvergedStepRunner(lambda s, d: me._step_into(s, d, None), fs, check_every, checks, multi, m)
            self._runners[key] = runner
        out, k, ratio = runner.run(fs, tol, max_steps)
        return self._out(out), k, ratio

    def _smooth(self, f, level, resid, ext, out=None):
        out = torch.empty_like(f) if out is None else out
        a = self._launch(f, out, level, nat.XB_MODE_SMOOTH, ext=ext, resid=resid)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step(smooth)")
        return out

    def _residual(self, f, level, resid, ext, coarse=None):
        """Restricted residual (coarse interior, SoA) and the workspace holding its
        squared-norm partials."""
        x0, x1, y0, y1 = self.interior_box(level)
        if coarse is None:
            coarse = torch.empty((8, (x1 - x0) // 2, (y1 - y0) // 2), dtype=f.dtype, device=f.device)
        lib = nat.lib()
        a = self._launch(f, None, level, nat.XB_MODE_RESIDUAL, ext=ext, resid=resid, coarse=coarse)
        n = int(lib.xb_elast_partials_count(a))
        ws = self._norm_ws.get(level)
        if ws is None or ws.partials.numel() < n:
            ws = self._norm_ws[level] = NormWorkspace(n, f.device)
        a.norm_partials = nat.ptr(ws.partials)
        nat.check(lib.xb_elast_step(a, nat.stream()), "xb_elast_step(residual)")
        return coarse, ws, n

    def _prolong_add(self, f, f_c, level):
        nx, ny = self.grid_size(level)
        cnx, cny = self.grid_size(level + 1)
        x0, x1, y0, y1 = self.interior_box(level)
        nat.check(nat.lib().xb_elast_prolong_add(nx, ny, self._code(f), nat.ptr(f), nat.ptr(f_c), cnx,
                                                 cny, x0, x1, y0, y1, nat.stream()), "xb_elast_prolong_add")

    def _add(self, a, b):
        out = torch.empty_like(a)
        nat.check(nat.lib().xb_lincomb(a.numel(), nat.dtype_code(a.dtype), 1.0, nat.ptr(a), 1.0, nat.ptr(b),
                                       nat.ptr(out), nat.stream()), "xb_lincomb")
        return out

    def _ext_soa(self, external_force, nonlinear_force, f, level):
        """``external_force (+ nonlinear_force(f))`` as a SoA buffer (elasticity.py:1259-1265)."""
        if not callable(nonlinear_force):
            return external_force
        fast = getattr(nonlinear_force, "soa_add", None)
        if fast is not None:
            return fast(f, external_force)
        nl = self._soa(nonlinear_force(view(f)), 2, level)
        return nl if external_force is None else self._add(external_force, nl)

    # ---- public API (elasticity.py:908-1344)
    def equilibrium(self, eta, sigma, cast_output=True, external_force=None):
        eta_s = self._soa(eta, 2)
        sig_s = self._soa(sigma, 3)
        ext = self._soa(external_force, 2)
        f = self._new(8)
        nat.check(nat.lib().xb_elast_equilibrium(self.nx, self.ny, nat.dtype_code(self._tdt), self.lamda, self.mu,
                                                 nat.ptr(eta_s), nat.ptr(sig_s), nat.ptr(self._force_soa),
                                                 nat.ptr(ext), nat.ptr(f), *self._fields(), nat.stream()),
                  "xb_elast_equilibrium")
        return self._out(f) if cast_output else view(f)

    def update_macroscopic(self, f, external_force=None):
        fs = self._soa(f, 8)
        ext = self._soa(external_force, 2)
        eta, sig = self._new(2), self._new(3)
        nat.check(nat.lib().xb_elast_macro(self.nx, self.ny, nat.dtype_code(self._tdt), self.lamda, self.mu,
                                           nat.ptr(fs), nat.ptr(self._force_soa), nat.ptr(ext), nat.ptr(eta),
                                           nat.ptr(sig), *self._fields(), nat.stream()), "xb_elast_macro")
        return view(eta), view(sig)

    def _divergence_soa(self, f, f_sub=None):
        nx, ny = f.shape[1], f.shape[2]
        out = torch.empty((1, nx, ny), dtype=f.dtype, device=f.device)
        fields = self._fields() if (nx, ny) == (self.nx, self.ny) else (None, None)
        nat.check(nat.lib().xb_elast_divergence(nx, ny, nat.dtype_code(f.dtype), self.lamda, self.mu, nat.ptr(f),
                                                nat.ptr(f_sub), nat.ptr(out), *fields, nat.stream()),
                  "xb_elast_divergence")
        return out

    def update_divergence(self, f):
        if isinstance(f, torch.Tensor) and f.dim() == 3 and f.shape[-1] == 8:
            nx, ny = f.shape[0], f.shape[1]
            return view(self._divergence_soa(as_soa(f, nx, ny, 8, self._tdt, self.device)))
        raise ValueError("update_divergence expects populations shaped (nx, ny, 8)")

    def collision(self, f, external_force, use_force=True):
        """Post-collision populations without boundary rules (elasticity.py:959-1000)."""
        fs = self._soa(f, 8)
        out, pc = torch.empty_like(fs), torch.empty_like(fs)
        a = self._launch(fs, out, 0, nat.XB_MODE_STEP, ext=self._soa(external_force, 2), fpc=pc,
                         use_force=use_force)
        a.bc = None
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step(collision)")
        return self._out(pc)

    def step(self, f_poststreaming, timestep, return_fpost=False, external_force=None, nonlinear_force=None):
        """elasticity.py:1128-1167."""
        if self.levels == 1:
            f = self._soa(f_poststreaming, 8)
            ext = self._ext_soa(self._soa(external_force, 2), nonlinear_force, f, 0)
            out, pc = self._run_step(f, ext, 0, return_fpost)
            return self._out(out), self._out(pc)
        return self._mg_step(f_poststreaming, timestep, external_force=external_force,
                             nonlinear_force=nonlinear_force)

    # ---- coarse levels run by the single-workgroup tail kernel
    def _tail_start(self):
        """First multigrid level run by the single-workgroup tail (None: no tail)."""
        if not getattr(self, "use_tail", True):
            return None
        for lvl in range(1, self.levels):
            nx, ny = self.grid_size(lvl)
            if nx * ny <= TAIL_CELLS:
                return lvl
        return None

    def _tail(self, level, nlf):
        """The tail that solves level ``level`` and below, if it starts there."""
        if self._tail_start() != level:
            return None
        # the plan holds lambda, mu, gamma, k_smooth and the coupling's scale by value
        # and the Darcy solver's gradient wall tables by address
        key = (level, None if nlf is None else (id(nlf.darcy), float(nlf.scale)), self.lamda, self.mu, self.gamma,
               self.k_smooth)
        tails = self.__dict__.setdefault("_tails", {})
        if key not in tails:
            tails[key] = VCycleTail(self, level, nlf)
        return tails[key]

    # ---- fused Biot coupling force: the sweep forms base + grad(div f) itself from a
    # carried divergence field (every sweep writes div of its output)
    @staticmethod
    def _fusable(nonlinear_force):
        return getattr(nonlinear_force, "fused", False)

    def _nl_args(self, a, div_in, div_out, nlf, level):
        nx, ny = self.grid_size(level)
        gt = nlf.darcy._grad_table(nx, ny)
        a.div_in = nat.ptr(div_in)
        a.div_out = nat.ptr(div_out)
        a.nl_scale = float(nlf.scale)
        a.grad_bc = gt.ref() if gt is not None else None
        return a

    def _step_nl(self, f, div, level, base, nlf, out=None, out_div=None):
        out = torch.empty_like(f) if out is None else out
        out_div = torch.empty_like(div) if out_div is None else out_div
        a = self._nl_args(self._launch(f, out, level, nat.XB_MODE_STEP, ext=base), div, out_div, nlf, level)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step(coupled)")
        return out, out_div

    def _smooth_nl(self, f, div, level, resid, base, nlf, out=None, out_div=None):
        out = torch.empty_like(f) if out is None else out
        out_div = torch.empty_like(div) if out_div is None else out_div
        a = self._nl_args(self._launch(f, out, level, nat.XB_MODE_SMOOTH, ext=base, resid=resid), div, out_div,
                          nlf, level)
        nat.check(nat.lib().xb_elast_step(a, nat.stream()), "xb_elast_step(coupled smooth)")
        return out, out_div

    def _residual_nl(self, f, div, level, resid, base, nlf, coarse=None):
        x0, x1, y0, y1 = self.interior_box(level)
        if coarse is None:
            coarse = torch.empty((8, (x1 - x0) // 2, (y1 - y0) // 2), dtype=f.dtype, device=f.device)
        lib = nat.lib()
        a = self._nl_args(self._launch(f, None, level, nat.XB_MODE_RESIDUAL, ext=base, resid=resid,
                                       coarse=coarse), div, None, nlf, level)
        n = int(lib.xb_elast_partials_count(a))
        ws = self._norm_ws.get(level)
        if ws is None or ws.partials.numel() < n:
            ws = self._norm_ws[level] = NormWorkspace(n, f.device)
        a.norm_partials = nat.ptr(ws.partials)
        nat.check(lib.xb_elast_step(a, nat.stream()), "xb_elast_step(coupled residual)")
        return coarse, ws, n

    def _coarse_solve_nl(self, f, div, level, residual, base, nlf):
        """Residual of ``f`` restricted to level ``level + 1`` and the coarse V-cycle
        from zero on it (the tail kernel where it starts there): ``(f_c, norm)``,
        norm (level 0 only) the residual's norm as a device scalar."""
        tail = self._tail(level + 1, nlf)
        res_c, ws, n = self._residual_nl(f, div, level, residual, base, nlf,
                                         coarse=tail.resid if tail is not None else None)
        norm = ws.finalize(n)[0] if level == 0 else None
        if tail is not None:
            return tail.launch(), norm
        # the coarse V-cycle from the zero state (None: no fill, no read)
        f_c, _ = self._vcycle_nl(None, None, level + 1, res_c, None, nlf)
        return f_c, norm

    def _smooth_tile_steps(self, level, nlf):
        """Smoothing steps per launch on LDS tiles at ``level`` (1: one sweep per
        step): ``self.smooth_tile_steps`` when set, else 2 on levels 0 and 1 and 4 below
        them (small grids whose launches are latency-bound: on the batched 64-instance
        set, with 8 smoothing steps, 4 physical steps took 0.786-0.792 s against
        0.805 s), where the coupling force's wall table fits the tile kernel's LDS
        table.  Every choice gives the same populations bit for bit."""
        m = getattr(self, "smooth_tile_steps", None)
        if m is None:
            m = 4 if level >= 2 else 2
        if m <= 1:
            return 1
        nx, ny = self.grid_size(level)
        if nlf is not None:
            gt = nlf.darcy._grad_table(nx, ny)
            if gt is not None and gt.nslots > nat.LDS_BC_SLOTS:
                return 1
        return min(m, 4 if nlf is not None else 6)

    def _smoothn_nl(self, f, div, level, resid, base, nlf, m, out=None, out_div=None, coarse=None):
        """``m`` coupled smoothing sweeps in one launch (``xb_elast_smoothn``), from
        ``f`` plus the injection of the next level's correction ``coarse`` if given;
        ``f`` None: from the zero state (``div`` is then not needed)."""
        nx, ny = self.grid_size(level)
        dt = self._tdt if f is None else f.dtype
        out = torch.empty((8, nx, ny), dtype=dt, device=self.device) if out is None else out
        out_div = torch.empty((1, nx, ny), dtype=dt, device=self.device) if out_div is None else out_div
        # with a zero start the divergence input is not read: out_div marks the coupling
        a = self._nl_args(self._launch(f, out, level, nat.XB_MODE_SMOOTH, ext=base, resid=resid),
                          div if f is not None else out_div, out_div, nlf, level)
        if coarse is not None:
            a.coarse_out = nat.ptr(coarse)
        nat.check(nat.lib().xb_elast_smoothn(a, int(m), nat.stream()), "xb_elast_smoothn")
        return out, out_div

    def _prolong_add_div(self, f, div, f_c, level):
        """``f += P f_c`` over the interior box in place, with ``div`` its divergence
        (reference src/elasticity.py:1290-1293)."""
        nxc, nyc = self.grid_size(level + 1)
        nx, ny = self.grid_size(level)
        x0, x1, y0, y1 = self.interior_box(level)
        nat.check(nat.lib().xb_elast_prolong_add_div(nx, ny, self._code(f), self.lamda, self.mu,
                                                     nat.ptr(f), nat.ptr(f_c), nxc, nyc, x0, x1, y0, y1,
                                                     nat.ptr(div), *(self._fields() if level == 0 else (None, None)),
                                                     nat.stream()), "xb_elast_prolong_add_div")

    def _smooth_seq_nl(self, f, div, level, residual, base, nlf, k, out=None, out_div=None, coarse=None):
        """``k`` coupled smoothing sweeps: tile launches of the tuned number of steps,
        the remainder as single sweeps; the last one writes ``out`` when given.
        ``coarse``: the next level's correction, prolonged into ``f`` first - inside
        the first tile launch when there is one (``f`` is then left as it was), else
        by the prolongation kernel in place.  ``f`` None: the zero state (read by no
        tile launch; zero-filled for single sweeps)."""
        m = self._smooth_tile_steps(level, nlf)
        if coarse is not None and not (m > 1 and k > 1):
            self._prolong_add_div(f, div, coarse, level)
            coarse = None
        if f is None and not (m > 1 and k > 1):
            # the zero start state, for the single sweeps
            nx, ny = self.grid_size(level)
            f = nat.fill(torch.empty((8, nx, ny), dtype=self._tdt, device=self.device), 0.0)
            div = nat.fill(torch.empty((1, nx, ny), dtype=self._tdt, device=self.device), 0.0)
        done = 0
        while done < k:
            step = min(m, k - done) if m > 1 else 1
            last = done + step == k
            o, od = (out, out_div) if last else (None, None)
            if step > 1:
                f, div = self._smoothn_nl(f, div, level, residual, base, nlf, step, o, od, coarse=coarse)
                coarse = None
            else:
                f, div = self._smooth_nl(f, div, level, residual, base, nlf, o, od)
            done += step
        return f, div

    def _vcycle_nl(self, f, div, level, residual, base, nlf, out=None, out_div=None, coarse=None):
        """``_vcycle`` with the Biot coupling force fused into the sweeps; ``div`` is
        the divergence of ``f`` and is carried along (``f`` None: the zero state of a
        coarse level).  Returns ``(f, div[, norm])``; ``coarse`` (a list) receives the
        coarse correction prolonged into ``f``."""
        f, div = self._smooth_seq_nl(f, div, level, residual, base, nlf, self.k_smooth)
        norm = None
        if level < self.levels - 1:
            f_c, norm = self._coarse_solve_nl(f, div, level, residual, base, nlf)
            if coarse is not None:
                coarse.append(f_c)
        else:
            f_c = None
        f, div = self._smooth_seq_nl(f, div, level, residual, base, nlf, self.k_smooth, out, out_div, coarse=f_c)
        if level == 0:
            return f, div, norm
        return f, div

    def _step_soa(self, f, external_force=None, nonlinear_force=None):
        """One ``step`` on SoA buffers: ``(f_new, residual_norm | None)``."""
        if self._fusable(nonlinear_force):
            div = self._divergence_soa(f)
            if self.levels == 1:
                out, _ = self._step_nl(f, div, 0, external_force, nonlinear_force)
                return out, None
            f, _, norm = self._vcycle_nl(f, div, 0, None, external_force, nonlinear_force)
            return f, norm
        if self.levels == 1:
            out, _ = self._run_step(f, self._ext_soa(external_force, nonlinear_force, f, 0), 0)
            return out, None
        return self._vcycle(f, 0, None, external_force, nonlinear_force)

    def _step(self, f_poststreaming, timestep, return_fpost=False, external_force=None, level=0):
        """elasticity.py:1169-1219."""
        f = self._soa(f_poststreaming, 8, level)
        out, pc = self._run_step(f, self._soa(external_force, 2, level), level, return_fpost)
        return self._out(out), self._out(pc)

    def _mg_step(self, f_poststreaming, timestep, level=0, residual=None, external_force=None, nonlinear_force=None):
        """elasticity.py:1221-1312; ``residual`` is shaped like the level's interior."""
        f = self._soa(f_poststreaming, 8, level)
        r = None
        if residual is not None:
            x0, x1, y0, y1 = self.interior_box(level)
            r = as_soa(residual, x1 - x0, y1 - y0, 8, self._tdt, self.device)
        out = self._vcycle(f, level, r, self._soa(external_force, 2, level), nonlinear_force)
        if level == 0:
            return self._out(out[0]), out[1]
        return self._out(out)

    def _vcycle(self, f, level, residual, external_force, nonlinear_force, out=None):
        """One V-cycle from ``level`` down (elasticity.py:1221-1312) on SoA buffers; the
        last smoothing step of the finest level writes into ``out`` when given."""
        for _ in range(self.k_smooth):
            ext = self._ext_soa(external_force, nonlinear_force, f, level)
            f = self._smooth(f, level, residual, ext)
        norm = None
        if level < self.levels - 1:
            ext = self._ext_soa(external_force, nonlinear_force, f, level)
            tail = self._tail(level + 1, None) if nonlinear_force is None else None
            res_c, ws, n = self._residual(f, level, residual, ext, coarse=tail.resid if tail is not None else None)
            if level == 0:
                norm = ws.finalize(n)[0]
            if tail is not None:
                f_c = tail.launch()
            else:
                nxc, nyc = self.grid_size(level + 1)
                f_c = nat.fill(torch.empty((8, nxc, nyc), dtype=f.dtype, device=f.device), 0.0)
                f_c = self._vcycle(f_c, level + 1, res_c, None, nonlinear_force)
            self._prolong_add(f, f_c, level)
        for k in range(self.k_smooth):
            ext = self._ext_soa(external_force, nonlinear_force, f, level)
            f = self._smooth(f, level, residual, ext, out if k == self.k_smooth - 1 else None)
        if level == 0:
            return f, norm
        return f

    def rhs_residual_norm(self, external_force=None, nonlinear_force=None):
        """Finest-level residual norm of the zero state, ``|S(0)|`` over the interior
        (the scale of the right-hand side), as a device scalar."""
        f0 = torch.zeros((8, self.nx, self.ny), dtype=self._tdt, device=self.device)
        if self._fusable(nonlinear_force):
            d0 = torch.zeros((1, self.nx, self.ny), dtype=self._tdt, device=self.device)
            _, ws, n = self._residual_nl(f0, d0, 0, None, external_force, nonlinear_force)
        else:
            ext = self._ext_soa(external_force, nonlinear_force, f0, 0)
            _, ws, n = self._residual(f0, 0, None, ext)
        return ws.finalize(n)[0]

    def vcycle_graph(self, external_force=None, nonlinear_force=None, keep_coarse=False):
        """A V-cycle captured as HIP graphs: see :class:`VCycleGraph`."""
        return VCycleGraph(self, external_force, nonlinear_force, keep_coarse)

    # ---- driver (elasticity.py:647-699, 1346-1456)
    def assign_fields_sharded(self):
        eta0, sigma0 = self.initialize_macroscopic_fields()
        if eta0 is None or sigma0 is None:
            return self._out(torch.zeros((8, self.nx, self.ny), dtype=self._tdt, device=self.device))
        return self.initialize_populations(eta0, sigma0)

    def initialize_populations(self, eta0, sigma0):
        return self.equilibrium(eta0, sigma0)

    def run(self, t_max):
        f = self.assign_fields_sharded()
        start = time.time()
        if self.ioRate == 0 and self.printInfoRate == 0 and not self.returnFpost and t_max >= 2:
            # nothing happens between steps: steps 0 and 1 eagerly (the MLUPS clock
            # starts after step 1), the remaining t_max - 1 from captured graphs
            for timestep in (0, 1):
                f, _ = self.step(f, timestep)
            torch.cuda.synchronize()
            start = time.time()
            if self.levels == 1:
                f = self.run_steps(f, t_max - 1)
            else:
                vg = self.vcycle_graph()
                vg.load(self._soa(f, 8))
                vg.run(t_max - 1)
                self.last_residual = vg.norm
                f = self._out(vg.state.clone())
            return self._finish_mlups(f, start, t_max)
        for timestep in range(0, t_max + 1):
            io_flag = self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max)
            print_iter_flag = self.printInfoRate > 0 and timestep % self.printInfoRate == 0
            if io_flag:
                eta_prev, sigma_prev = (io_field(x, self.downsamplingFactor) for x in self.update_macroscopic(f))
            f, fstar = self.step(f, timestep, return_fpost=self.returnFpost)
            if print_iter_flag:
                print(f"Timestep {timestep} of {t_max} completed")
            if io_flag:
                eta, sigma = (io_field(x, self.downsamplingFactor) for x in self.update_macroscopic(f))
                self.handle_io_timestep(timestep, f, fstar, eta, sigma, eta_prev, sigma_prev)
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

    def handle_io_timestep(self, timestep, f, fstar, eta, sigma, eta_prev, sigma_prev):
        self.output_data(timestep=timestep, eta=eta, eta_prev=eta_prev, sigma=sigma, sigma_prev=sigma_prev,
                         f_poststreaming=f, f_postcollision=fstar)


# Multigrid levels of at most this many cells run as one single-workgroup kernel
# (xb_vcycle_tail): their sweeps are bound by launch and load latency.  Measured on
# MI355X (V-cycle time, threshold 0 / 512 / 2048 / 8192 cells): 1024 x 1026 with 10
# levels FP32 3.10 / 2.59 / 2.72 / 3.61 ms, 256^2 with 8 levels FP64 298 / 238 / 243
# / 371 us; above about a thousand cells one workgroup is slower than a launch.
TAIL_CELLS = 512


class VCycleTail:
    """The levels ``start`` .. of the V-cycles of ``el`` (below the level that calls
    them) as one single-workgroup launch from the zero state (``xb_vcycle_tail``):
    bit-identical to the level-by-level launches.  The restricted residual of level
    ``start - 1`` is written into :attr:`resid`; the result is :attr:`result`."""

    def __init__(self, el, start, nlf):
        no_capture("VCycleTail")  # its device table is copied from host memory
        lib = nat.lib()
        coupled = nlf is not None
        self.levels = list(range(start, el.levels))
        n = len(self.levels)
        dt, dev = el._tdt, el.device
        arr = (nat.TailLevel * n)()
        self._keep = []
        # the rule tables whose device addresses the plan's table holds
        self._tables = []
        for i, lvl in enumerate(self.levels):
            nx, ny = el.grid_size(lvl)
            x0, x1, y0, y1 = el.interior_box(lvl)
            b = {"f0": torch.zeros((8, nx, ny), dtype=dt, device=dev),
                 "f1": torch.zeros((8, nx, ny), dtype=dt, device=dev),
                 "resid": torch.zeros((8, x1 - x0, y1 - y0), dtype=dt, device=dev),
                 "scratch": torch.zeros((8, nx, ny), dtype=dt, device=dev)}
            if coupled:
                b["div0"] = torch.zeros((1, nx, ny), dtype=dt, device=dev)
                b["div1"] = torch.zeros((1, nx, ny), dtype=dt, device=dev)
            table = el._tables.get(lvl)
            gt = nlf.darcy._grad_table(nx, ny) if coupled else None
            self._tables += [table, gt]
            arr[i] = nat.TailLevel(nx, ny, x0, x1, y0, y1, table.ref() if table is not None else None,
                                   gt.ref() if gt is not None else None, nat.ptr(b["f0"]), nat.ptr(b["f1"]),
                                   nat.ptr(b.get("div0")), nat.ptr(b.get("div1")), nat.ptr(b["resid"]),
                                   nat.ptr(b["scratch"]))
            self._keep.append(b)
        nbytes = int(lib.xb_vcycle_tail_bytes(n))
        self.table = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        self.plan = nat.TailPlan()
        scale = float(nlf.scale) if coupled else 0.0
        nat.check(lib.xb_vcycle_tail_prepare(nat.dtype_code(dt), el.lamda, el.mu, el.gamma, el.k_smooth, scale,
                                             int(coupled), arr, n, nat.ptr(self.table), nbytes,
                                             ctypes.byref(self.plan)), "xb_vcycle_tail_prepare")
        if not getattr(el, "tail_in_lds", True):
            # the global-memory form even where the levels fit the LDS arena
            self.plan.lds_bytes = 0
        self.resid = self._keep[0]["resid"]
        self.result = self._keep[0]["f0"]

    def launch(self):
        nat.check(nat.lib().xb_vcycle_tail(ctypes.byref(self.plan), nat.stream()), "xb_vcycle_tail")
        return self.result


class VCycleGraph:
    """V-cycles replayed from two captured HIP graphs (ping-pong state buffers).

    The external force is a static buffer: :meth:`set_force` copies a new force in
    before the next replays (the Biot coupler changes it once per physical step).
    Each replay advances the state by one V-cycle and leaves the level-0 residual
    norm of that cycle in :attr:`norm` (a device scalar; reading it synchronises)."""

    def __init__(self, el, external_force=None, nonlinear_force=None, keep_coarse=False):
        if el.levels < 2:
            raise ValueError("V-cycle graphs need MG_levels >= 2")
        self.el = el
        shape = (8, el.nx, el.ny)
        self.buf = [torch.zeros(shape, dtype=el._tdt, device=el.device) for _ in range(2)]
        self.ext = None if external_force is None else external_force.clone()
        nlf = nonlinear_force
        # the captured kernels read the coupling's Darcy gradient wall tables and the
        # coarse tails by address: hold their owners as long as the graphs live
        self.nlf = nlf
        self.fused = el._fusable(nlf)
        self.div = [torch.zeros((1, el.nx, el.ny), dtype=el._tdt, device=el.device) for _ in range(2)]

        if keep_coarse and not self.fused:
            raise ValueError("keep_coarse needs the fused coupling force")

        def cycle(src, coarse=None):
            if self.fused:
                _, _, norm = el._vcycle_nl(self.buf[src], self.div[src], 0, None, self.ext, nlf,
                                           out=self.buf[1 - src], out_div=self.div[1 - src], coarse=coarse)
            else:
                _, norm = el._vcycle(self.buf[src], 0, None, self.ext, nlf, out=self.buf[1 - src])
            return norm
        # eager warm-up: rule tables, norm workspaces, gradient tables and coarse tails
        # exist before capture
        cycle(0)
        torch.cuda.synchronize()
        self.tails = dict(el.__dict__.get("_tails", {}))
        self.graphs, self.norms, self.coarse_bufs = [], [], []
        for src in (0, 1):
            g = torch.cuda.CUDAGraph()
            kept = [] if keep_coarse else None
            with graph_capture(g):
                norm = cycle(src, kept)
            self.graphs.append(g)
            self.norms.append(norm)
            # the graph's coarse correction buffer (held here, so the graph's memory pool
            # never hands it to another allocation of the capture)
            self.coarse_bufs.append(kept[0] if keep_coarse else None)
        self.cur = 0
        self.norm = None
        self.coarse = None

    @property
    def state(self):
        """Current populations (SoA buffer owned by the graph)."""
        return self.buf[self.cur]

    def load(self, f_soa):
        self.buf[self.cur].copy_(f_soa)
        if self.fused:
            el = self.el
            nat.check(nat.lib().xb_elast_divergence(el.nx, el.ny, nat.dtype_code(f_soa.dtype), el.lamda, el.mu,
                                                    nat.ptr(self.buf[self.cur]), None, nat.ptr(self.div[self.cur]),
                                                    *el._fields(), nat.stream()), "xb_elast_divergence")

    def set_force(self, force_soa):
        if self.ext is None:
            raise ValueError("this V-cycle graph was captured without an external force")
        self.ext.copy_(force_soa)

    def replay(self):
        g = self.graphs[self.cur]
        self.norm = self.norms[self.cur]
        self.coarse = self.coarse_bufs[self.cur]  # with keep_coarse: this cycle's coarse correction
        g.replay()
        self.cur = 1 - self.cur
        return self.norm

    def run(self, n):
        for _ in range(n):
            self.replay()

    def run_to(self, tol, max_cycles, patience=None, r_ref=None):
        """Replay until the cycle's residual norm is below ``tol`` times the first
        cycle's, or times ``r_ref`` when that is given and positive (one
        device-to-host read per cycle).  With ``patience``, also stop once the residual has not improved for
        that many cycles, restoring the state of the best cycle.  Returns ``(cycles,
        best relative residual, reached)``; :attr:`last_r0` is the first cycle's norm."""
        r0 = None
        scale = None
        best, best_k, best_state = None, 0, None
        for k in range(max_cycles):
            self.replay()
            r = self.norm.item()
            if not np.isfinite(r):
                raise FloatingPointError(f"V-cycle residual is not finite after {k + 1} cycles")
            if r0 is None:
                r0 = self.last_r0 = r
                scale = r_ref if r_ref is not None and r_ref > 0.0 else r0
            if r <= tol * scale:
                return k + 1, (r / scale if scale > 0 else 0.0), True
            if patience is not None:
                if best is None or r < best:
                    best, best_k = r, k
                    # the norm is that of the state entering this cycle: keep that state
                    best_state = self.buf[1 - self.cur].clone() if k > 0 else None
                elif k - best_k >= patience:
                    if best_state is not None:
                        self.buf[self.cur].copy_(best_state)
                    return k + 1, best / scale, False
        return max_cycles, (best if best is not None else r) / scale, False


# ---------------------------------------------------------------------------
# Elastic boundary conditions (elasticity.py:1503-1750).  Construction follows the
# reference; ``apply_rules`` is the symbolic form of each ``apply``.

class ElasticityBounceBackFullway(BoundaryCondition):
    """Full-way bounce-back (Dirichlet), post-collision at the wall cells:
    ``f*_i = f_opp(i) + 6 w_i c_i . eta`` (elasticity.py:1503-1550)."""

    def __init__(self, indices, gridInfo, precision_policy, eta=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityBounceBackFullway"
        self.implementationStep = "PostCollision"
        self.eta = None if eta is None else np.asarray(eta, dtype=np.float64)

    def apply_rules(self, state):
        opp = np.asarray(self.lattice.opp_indices, dtype=np.int64)
        state.code[:, :] = opp[None, :] + 1
        state.const[:, :] = 0.0
        if self.eta is not None:
            state.const[:, :] = np.broadcast_to(self.cu_rules(self.eta), state.const.shape)


class ElasticityBounceBack(BoundaryCondition):
    """Halfway bounce-back (Dirichlet), post-streaming (elasticity.py:1552-1646)."""

    def __init__(self, indices, gridInfo, precision_policy, eta=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.eta = None if eta is None else np.asarray(eta, dtype=np.float64)

    def configure(self, boundaryMask):
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.eta is not None:
            if self.eta.shape == (self.lattice.d,):
                self.eta = np.tile(self.eta, (nbd_modified, 1))
            elif nbd_orig != nbd_modified:
                print("WARNING: Assuming a correct displacement vector to impose at the BC cells!")
                self.eta = self.eta[array_indices]

    def apply_rules(self, state):
        state.set_from_fin(self.imissing, self.iknown, +1)
        if self.eta is not None:
            cu = np.broadcast_to(self.cu_rules(self.eta), state.const.shape)
            im = np.asarray(self.imissing, dtype=np.int64)
            state.add(im, np.take_along_axis(cu, im, 1))


class ElasticityAntiBounceBack(BoundaryCondition):
    """Halfway anti-bounce-back (Neumann / traction), post-streaming
    (elasticity.py:1648-1750).  Missing and known directions pair in row-major mask order."""

    def __init__(self, indices, gridInfo, precision_policy, T=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityAntiBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.T = None if T is None else np.asarray(T, dtype=np.float64)

    def configure(self, boundaryMask):
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.T is not None:
            if self.T.shape == (self.lattice.d,):
                self.T = np.tile(self.T, (nbd_modified, 1))
            elif nbd_orig != nbd_modified:
                print("WARNING: Assuming a correct normal stress vector to impose at the BC cells!")
                self.T = self.T[array_indices]

    def get_missing_mask(self, boundaryMask):
        if self.T is not None:
            normals = self.get_normals(boundaryMask)
            T = self.T
            if np.all(normals[:, 0] == -1):  # left
                self.i_update = (0, 4, 7)
                self.T_update = np.stack((T[:, 0], 0.5 * T[:, 1], -0.5 * T[:, 1]), axis=-1)
            elif np.all(normals[:, 0] == 1):  # right
                self.i_update = (2, 5, 6)
                self.T_update = np.stack((-T[:, 0], 0.5 * T[:, 1], -0.5 * T[:, 1]), axis=-1)
            elif np.all(normals[:, 1] == -1):  # bottom
                self.i_update = (1, 4, 5)
                self.T_update = np.stack((T[:, 1], 0.5 * T[:, 0], -0.5 * T[:, 0]), axis=-1)
            elif np.all(normals[:, 1] == 1):  # top
                self.i_update = (3, 6, 7)
                self.T_update = np.stack((-T[:, 1], -0.5 * T[:, 0], 0.5 * T[:, 0]), axis=-1)
            else:
                raise NotImplementedError
        return super().get_missing_mask(boundaryMask)

    def apply_rules(self, state):
        state.set_masked_from_fin(self.imissingMask, self.iknownMask, -1)
        if self.T is not None:
            cols = np.broadcast_to(np.asarray(self.i_update, dtype=np.int64), self.T_update.shape)
            state.add(cols, self.T_update)
