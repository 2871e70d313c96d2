"""Shared host machinery of the solvers: precision policy, field layout, LBMBase.

Field layout.  Populations and fields live on the GPU as contiguous
structure-of-arrays tensors ``(c, nx, ny)`` (y fastest).  What the API hands out
and accepts are the reference-shaped ``(nx, ny, c)`` views of those buffers
(``soa.permute(1, 2, 0)``), so user code indexes ``f[x, y, i]`` as with the
reference's arrays, and passing a returned view back costs no copy.

``LBMBase`` keeps the parts of the reference ``src/base.py`` that the Darcy
solver uses (construction :44-146, boundary data :351-372,
``distributed_array_init`` :391-411, grid mask :413-449, bounding box
:451-485, precisions :487-510, ``assign_fields_sharded`` :528-559).
"""
import contextlib
import ctypes
import gc
import time

import numpy as np
import torch

from . import _native as nat
from .boundary_conditions import configure_bcs, create_grid_mask, build_rule_table

_NP2TORCH = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
             np.dtype(np.float16): torch.float16, np.dtype(np.bool_): torch.bool,
             np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64}


def torch_dtype(dt):
    if isinstance(dt, torch.dtype):
        return dt
    return _NP2TORCH[np.dtype(dt)]


class PrecisionPolicy:
    """``"compute/storage"`` precision (stand-in for the reference's ``jmp.Policy``).

    ``compute_dtype`` / ``output_dtype`` are NumPy dtypes, so driver hooks that pass
    them to NumPy keep working; ``torch_compute`` / ``torch_output`` are the torch ones.
    Supported: ``"f64/f64"``, ``"f32/f32"``, and the narrower-storage policies
    ``"f64/f32"``, ``"f64/f16"`` and ``"f32/f16"`` (``mixed``: buffers and arithmetic in
    the compute precision, every population value rounded to the storage precision
    where the reference stores it, the states handed out in the storage precision;
    ``storage_code`` is their kernel dtype).  ``"f16/f16"`` (FP16 arithmetic) raises.
    """
    TABLE = {"f64/f64": (np.float64, np.float64), "f32/f32": (np.float32, np.float32),
             "f32/f16": (np.float32, np.float16), "f16/f16": (np.float16, np.float16),
             "f64/f32": (np.float64, np.float32), "f64/f16": (np.float64, np.float16)}

    def __init__(self, precision):
        # elasticity.py:606-629: unknown strings fall back to f32/f32
        cdt, odt = self.TABLE.get(precision, (np.float32, np.float32))
        if cdt is np.float16:
            raise NotImplementedError(
                f"precision {precision!r}: half-precision arithmetic is not implemented on this backend "
                "(FP16 storage: 'f32/f16' or 'f64/f16')")
        self.compute_dtype = cdt
        self.param_dtype = cdt
        self.output_dtype = odt
        self.torch_compute = torch_dtype(cdt)
        self.torch_output = torch_dtype(odt)
        self.mixed = cdt is not odt
        self.storage_code = {(np.float64, np.float32): nat.XB_F64_S32, (np.float64, np.float16): nat.XB_F64_S16,
                             (np.float32, np.float16): nat.XB_F32_S16}.get((cdt, odt))

    def require_single(self, what):
        """Raise for the narrower-storage policies in paths that have no such form."""
        if self.mixed:
            raise NotImplementedError(f"{what} is implemented for 'f64/f64' and 'f32/f32' only")

    def cast_to_compute(self, x):
        return None if x is None else x.to(self.torch_compute)

    def cast_to_output(self, x):
        return None if x is None else x.to(self.torch_output)


def default_device():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


def as_soa(x, nx, ny, c, dtype, device):
    """``(nx, ny, c)`` array-like -> contiguous ``(c, nx, ny)`` tensor on ``device``.

    A view returned by this package (the permuted SoA buffer) is passed through
    without a copy."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.dim() == 2:
            t = t.unsqueeze(-1)
        if tuple(t.shape) == (nx, ny, c) and t.device == device and t.dtype == dtype:
            b = t.permute(2, 0, 1)
            if b.is_contiguous():
                return b
        t = t.to(device=device, dtype=dtype)
    else:
        t = torch.as_tensor(np.asarray(x), dtype=dtype, device=device)
        if t.dim() == 2:
            t = t.unsqueeze(-1)
    if tuple(t.shape) != (nx, ny, c):
        t = t.expand(nx, ny, c)
    return t.permute(2, 0, 1).contiguous()


def view(soa):
    """Reference-shaped ``(nx, ny, c)`` view of a ``(c, nx, ny)`` buffer."""
    return soa.permute(1, 2, 0)


def io_field(x, factor=1):
    """A macroscopic field as the run loops hand it to ``output_data``: down-sampled by
    ``factor`` and gathered to host memory with a leading process axis, as
    ``process_allgather`` returns it (one process: ``(1, nx, ny, c)``; elasticity.py:
    1378-1401, darcy.py:200-229, poroelasticity.py:273-277); drivers read
    ``np.array(kwargs["eta"])[0]``."""
    from .utils import downsample_field
    x = downsample_field(x, factor)
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.asarray(x)[None]


def capturing():
    """True while the current stream records a HIP graph."""
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


def no_capture(what):
    """Raise if called while a graph is being captured (for set-up work that copies
    host data synchronously or allocates outside the capture's memory pool)."""
    if capturing():
        raise RuntimeError(f"{what} must be created before a graph capture, not inside one")


@contextlib.contextmanager
def graph_capture(graph):
    """``torch.cuda.graph(graph)`` with Python's cyclic garbage collector held off.

    Solvers cache runners, adjoint sessions and V-cycle graphs that refer back to
    them; such reference cycles are freed only by the cyclic collector, whenever an
    allocation happens to trigger it.  Run inside a capture, that collection would
    destroy ``CUDAGraph`` objects (HIP graph executables, their memory pools and
    the device buffers they own) while a global-mode stream capture is in
    progress, which the HIP runtime does not support.  So pending cycles are
    collected first, outside the capture, and the collector stays disabled until
    the capture has ended."""
    if capturing():
        raise RuntimeError("graph captures do not nest")
    gc.collect()
    torch.cuda.synchronize()
    was_enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.graph(graph):
            yield graph
    finally:
        if was_enabled:
            gc.enable()


class GraphedSteps:
    """``n`` (even) ping-pong steps ``launch(src, dst)`` captured once into a HIP graph
    (through ``torch.cuda.CUDAGraph``) and replayed: one launch for the whole chunk.
    The state lives in ``buf_a`` before and after each replay."""

    def __init__(self, launch, buf_a, buf_b, n):
        if n < 2 or n % 2:
            raise ValueError("graphed chunks hold an even number of steps")
        self.n = n
        self.buf_a, self.buf_b = buf_a, buf_b
        self.graph = torch.cuda.CUDAGraph()
        with graph_capture(self.graph):
            for _ in range(n // 2):
                launch(buf_a, buf_b)
                launch(buf_b, buf_a)

    def replay(self):
        self.graph.replay()


class StepRunner:
    """Persistent ping-pong buffers and one captured graph of about ``chunk`` steps;
    :meth:`run` advances a state by any number of steps (graph replays, then
    multi-step launches, then single launches) and returns a fresh copy of the
    result.  ``launch(src, dst)`` is one step; ``multi(src, dst)``, when given, does
    ``m`` steps in one launch, and the graph then holds an even number of those."""

    def __init__(self, launch, like, chunk, multi=None, m=1):
        if multi is None:
            m = 1
        self.launch, self.multi, self.m = launch, multi, m
        self.a = torch.empty_like(like)
        self.b = torch.empty_like(like)
        self.a.copy_(like)
        launch(self.a, self.b)  # warm-up outside the capture
        if multi is not None:
            multi(self.a, self.b)
        torch.cuda.synchronize()
        step = multi if multi is not None else launch
        n = 2 * (chunk // (2 * m))
        self.graph = GraphedSteps(step, self.a, self.b, n) if n >= 2 else None
        self.gsteps = n * m

    def run(self, f, n):
        self.a.copy_(f)
        done = 0
        if self.graph is not None:
            while n - done >= self.gsteps:
                self.graph.replay()
                done += self.gsteps
        if self.multi is not None:
            while n - done >= self.m:
                self.multi(self.a, self.b)
                self.a, self.b = self.b, self.a
                done += self.m
        while done < n:
            self.launch(self.a, self.b)
            self.a, self.b = self.b, self.a
            done += 1
        if self.graph is not None and self.a is not self.graph.buf_a:
            self.a, self.b = self.b, self.a
            self.a.copy_(self.b)
        return self.a.clone()


class ConvergedStepRunner:
    """Step loop to a relative step-to-step residual, tested on the device.

    From ``f_0`` it takes one step, records ``r_0 = |f_1 - f_0|`` and then, in blocks of
    ``check_every`` steps, ``r_k = |f_k - f_{k-1}|`` at ``k = 1 + j check_every``; the
    first ``k`` with ``r_k < tol r_0`` (or a non-finite ``r_k``) latches ``k``, ``r_k / r_0``
    and ``f_k`` into device buffers.  ``checks`` blocks are one captured graph; the host
    replays it and reads the latch once per replay, so a solve costs one
    synchronisation per ``checks * check_every`` steps and runs at most that many
    steps past the latched one (their results are discarded).  The norms are the
    library's FP64 partial sums (``xb_sqdiff_partials`` + ``xb_norm_finalize``).

    ``launch(src, dst)`` is one step, ``multi(src, dst)`` (optional) ``m`` steps."""

    def __init__(self, launch, like, check_every=10, checks=10, multi=None, m=1):
        if check_every < 2 or checks < 1:
            raise ValueError("check_every must be >= 2 and checks >= 1")
        no_capture("ConvergedStepRunner")
        if multi is None:
            m = 1
        self.launch, self.multi, self.m = launch, multi, m
        self.ce, self.checks = int(check_every), int(checks)
        dev = like.device
        self.a = torch.empty_like(like)
        self.b = torch.empty_like(like)
        self.out = torch.empty_like(like)
        self.n = like.numel()
        self.dtype_code = nat.dtype_code(like.dtype)
        self.partials = torch.zeros(2048, dtype=torch.float64, device=dev)
        self.count = nat.c_int32(0)
        self.rv = torch.zeros(2, dtype=torch.float64, device=dev)
        self.r0 = torch.zeros(2, dtype=torch.float64, device=dev)
        self.tol = torch.zeros((), dtype=torch.float64, device=dev)
        self.k = torch.zeros((), dtype=torch.int64, device=dev)
        self.done = torch.full((), -1, dtype=torch.int64, device=dev)
        self.rel = torch.zeros((), dtype=torch.float64, device=dev)
        self.a.copy_(like)
        self._block()  # warm-up outside the capture (buffers, kernels)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with graph_capture(self.graph):
            for _ in range(self.checks):
                self._block()

    def _norm(self, x, y, out):
        lib = nat.lib()
        nat.check(lib.xb_sqdiff_partials(self.n, self.dtype_code, nat.ptr(x), nat.ptr(y), nat.ptr(self.partials),
                                         2048, ctypes.byref(self.count), nat.stream()), "xb_sqdiff_partials")
        nat.check(lib.xb_norm_finalize(nat.ptr(self.partials), int(self.count.value), nat.ptr(out), nat.stream()),
                  "xb_norm_finalize")

    def _advance(self, n):
        """``n`` steps from buffer a; returns the buffer holding the result."""
        src, dst = self.a, self.b
        done = 0
        if self.multi is not None:
            while n - done >= self.m:
                self.multi(src, dst)
                src, dst = dst, src
                done += self.m
        while done < n:
            self.launch(src, dst)
            src, dst = dst, src
            done += 1
        return src

    def _block(self):
        """``check_every`` steps from buffer a back into buffer a, the last one tested."""
        x = self._advance(self.ce - 1)
        y = self.b if x is self.a else self.a
        self.launch(x, y)
        self._norm(x, y, self.rv)
        self.k.add_(self.ce)
        hit = ((self.rv[0] / self.r0[0] < self.tol) | (self.rv[1] != 0)) & (self.done < 0)
        self.done.copy_(torch.where(hit, self.k, self.done))
        self.rel.copy_(torch.where(hit, self.rv[0] / self.r0[0], self.rel))
        self.out.copy_(torch.where(hit, y, self.out))
        if y is not self.a:
            nat.check(nat.lib().xb_lincomb(self.n, self.dtype_code, 1.0, nat.ptr(y), 0.0, nat.ptr(y), nat.ptr(self.a),
                                           nat.stream()), "xb_lincomb")

    def run(self, f, tol, max_steps):
        """``(f_k, k, r_k / r_0)``; ``k`` is -1 when ``max_steps`` passed first."""
        self.a.copy_(f)
        self.launch(self.a, self.b)
        self._norm(self.a, self.b, self.r0)
        self.a.copy_(self.b)
        self.tol.fill_(float(tol))
        self.k.fill_(1)
        self.done.fill_(-1)
        self.out.copy_(self.a)
        self.rel.fill_(1.0)
        steps = 1
        while steps < max_steps:
            self.graph.replay()
            steps += self.checks * self.ce
            if int(self.done.item()) >= 0:
                break
        k = int(self.done.item())
        if k < 0:
            return self.a.clone(), -1, float(self.rv[0].item() / self.r0[0].item())
        return self.out.clone(), k, float(self.rel.item())


class NormWorkspace:
    """FP64 partial sums + the finalised (norm, non-finite flag) pair."""

    def __init__(self, n, device):
        no_capture("NormWorkspace")
        self.partials = torch.zeros(max(int(n), 1), dtype=torch.float64, device=device)
        self.count = nat.c_int32(0)

    def finalize(self, n):
        out = torch.empty(2, dtype=torch.float64, device=self.partials.device)
        nat.check(nat.lib().xb_norm_finalize(nat.ptr(self.partials), int(n), nat.ptr(out), nat.stream()),
                  "xb_norm_finalize")
        return out


class LBMBase(object):
    """Base of the D2Q9 Darcy solver (reference src/base.py)."""

    def __init__(self, **kwargs):
        self.omega = kwargs.get("omega")
        self.nz = kwargs.get("nz")
        self.nx = kwargs.get("nx")
        self.ny = kwargs.get("ny")
        self.precision = kwargs.get("precision")
        self.precisionPolicy = PrecisionPolicy(self.precision)
        self.lattice = kwargs.get("lattice")
        self.checkpointRate = kwargs.get("checkpoint_rate", 0)
        self.checkpointDir = kwargs.get("checkpoint_dir", "./checkpoints")
        self.downsamplingFactor = kwargs.get("downsampling_factor", 1)
        self.printInfoRate = kwargs.get("print_info_rate", 100)
        self.ioRate = kwargs.get("io_rate", 0)
        self.returnFpost = kwargs.get("return_fpost", False)
        self.computeMLUPS = kwargs.get("compute_MLUPS", False)
        self.restore_checkpoint = kwargs.get("restore_checkpoint", False)
        self.nDevices = 1
        self.backend = "rocm"
        self.device = default_device()
        self.sharding = None
        if self.computeMLUPS:
            self.restore_checkpoint = False
            self.ioRate = 0
            self.checkpointRate = 0
            self.printInfoRate = 0
        self.c = self.lattice.c
        self.q = self.lattice.q
        self.w = self.lattice.w
        self.dim = self.lattice.d
        if None in {self.nx, self.ny, self.nz}:
            raise ValueError("nx, ny, and nz must be provided. For 2D examples, nz must be set to 0.")
        self.gridInfo = {"nx": self.nx, "ny": self.ny, "nz": self.nz, "dim": self.lattice.d, "lattice": self.lattice}
        self.boundingBoxIndices = self.bounding_box_indices()
        self._create_boundary_data()
        self.force = self.get_force()

    # ---- validated attributes (base.py:148-320)
    @property
    def lattice(self):
        return self._lattice

    @lattice.setter
    def lattice(self, value):
        if value is None:
            raise ValueError("Lattice type must be provided.")
        if self.nz == 0 and value.name not in ["D2Q9"]:
            raise ValueError("For 2D simulations, lattice type must be LatticeD2Q9.")
        if self.nz != 0:
            raise ValueError("3D simulations are not supported on this path.")
        self._lattice = value

    @property
    def omega(self):
        return self._omega

    @omega.setter
    def omega(self, value):
        if value is None:
            raise ValueError("omega must be provided")
        if not isinstance(value, float):
            raise TypeError("omega must be a float")
        self._omega = value

    def _int_attr(name):
        def getter(self):
            return getattr(self, "_" + name)

        def setter(self, value):
            if value is None:
                raise ValueError(f"{name} must be provided")
            if not isinstance(value, int) or isinstance(value, bool):
                raise TypeError(f"{name} must be an integer")
            setattr(self, "_" + name, value)
        return property(getter, setter)

    nx = _int_attr("nx")
    ny = _int_attr("ny")
    nz = _int_attr("nz")

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, value):
        if not isinstance(value, str):
            raise TypeError("precision must be a string")
        self._precision = value

    def set_precisions(self, precision):
        return PrecisionPolicy.TABLE.get(precision, (np.float32, np.float32))

    # ---- geometry and boundary data
    def bounding_box_indices(self):
        nx, ny = self.nx, self.ny
        return {"bottom": np.array([[i, 0] for i in range(nx)], dtype=int),
                "top": np.array([[i, ny - 1] for i in range(nx)], dtype=int),
                "left": np.array([[0, i] for i in range(ny)], dtype=int),
                "right": np.array([[nx - 1, i] for i in range(ny)], dtype=int)}

    def create_grid_mask(self, solid_halo_voxels):
        return create_grid_mask(self.nx, self.ny, self.lattice, solid_halo_voxels)

    def _create_boundary_data(self):
        self.BCs = []
        self.set_boundary_conditions()
        start = time.time()
        configure_bcs(self.BCs, self.nx, self.ny, self.lattice)
        self._bc_table = build_rule_table(self.BCs, self.nx, self.ny, self.q, self.precisionPolicy.torch_compute,
                                          self.device)
        self._bc_build_time = time.time() - start

    def distributed_array_init(self, shape, type, init_val=0, sharding=None):
        """``jnp.full(shape, init_val, dtype=type)`` on the solver's device (base.py:391-411)."""
        dt = torch_dtype(type)
        if isinstance(init_val, torch.Tensor):
            v = init_val.to(device=self.device, dtype=dt)
        else:
            v = torch.as_tensor(np.asarray(init_val), dtype=dt, device=self.device)
        return v.expand(shape).clone() if tuple(v.shape) != tuple(shape) else v.clone()

    # ---- hooks
    def set_boundary_conditions(self):
        pass

    def get_force(self):
        pass

    def initialize_macroscopic_fields(self):
        print("WARNING: Default initial conditions assumed: density = 1, velocity = 0")
        print("         To set explicit initial density and velocity, use self.initialize_macroscopic_fields.")
        return None, None

    def output_data(self, **kwargs):
        pass

    def show_simulation_parameters(self):
        names = ["omega", "nx", "ny", "nz", "dim", "precision", "lattice", "ioRate", "computeMLUPS", "backend",
                 "nDevices"]
        print(f"**** Simulation Parameters for {self.__class__.__name__} ****")
        for n in names:
            print(f"{n:>30} | {getattr(self, n, 'Attribute not set')}")
