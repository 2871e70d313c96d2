"""D2Q9 BGK advection-diffusion-reaction solver used as the Darcy solver.  Oracle only.

Restates ``src/darcy.py`` (``AdvectionDiffusionReactionBGK``) and the parts of
``src/base.py`` it uses (construction, grid mask, ``assign_fields_sharded``,
streaming, ``apply_bc``), on torch-CPU arrays of layout ``(nx, ny, 9)``.

Extension (not in the reference): ``omega`` may be a per-cell field of shape
``(nx, ny, 1)``; a scalar ``omega`` reproduces the reference exactly.
"""
import numpy as np
import torch

from . import bc as obc
from .bc import PrecisionPolicy, gathered, stream


class OracleDarcy:
    def __init__(self, **kwargs):
        # base.py:44-146, darcy.py:30-34
        self.omega = kwargs.get("omega")
        self.nx = kwargs.get("nx")
        self.ny = kwargs.get("ny")
        self.nz = kwargs.get("nz", 0)
        self.precision = kwargs.get("precision", "f64/f64")
        self.precisionPolicy = PrecisionPolicy(self.precision)
        self.lattice = kwargs.get("lattice")
        self.returnFpost = kwargs.get("return_fpost", False)
        self.ioRate = kwargs.get("io_rate", 0)
        self.c = self.lattice.c
        self.q = self.lattice.q
        self.w = self.lattice.w
        self.dim = 2
        self.gridInfo = {"nx": self.nx, "ny": self.ny, "nz": 0, "dim": 2, "lattice": self.lattice}
        self.boundingBoxIndices = self.bounding_box_indices()
        self.BCs = []
        self.set_boundary_conditions()
        obc.build_bcs(self.BCs, self.nx, self.ny, self.lattice)
        self.vel = kwargs.get("vel", None)
        if self.vel is not None:
            raise NotImplementedError("advection velocity is not on the Darcy path")
        src = self.get_source()
        self.source = None if src is None else torch.as_tensor(np.asarray(src), dtype=self.precisionPolicy.compute_dtype)

    def _omega(self, dtype):
        if isinstance(self.omega, float):
            return self.omega
        if isinstance(self.omega, torch.Tensor):
            return self.omega.to(dtype)  # keeps an autograd graph (gradient checks)
        return torch.as_tensor(np.asarray(self.omega), dtype=dtype)

    def bounding_box_indices(self):
        nx, ny = self.nx, self.ny
        return {"bottom": np.array([[i, 0] for i in range(nx)], dtype=int),
                "top": np.array([[i, ny - 1] for i in range(nx)], dtype=int),
                "left": np.array([[0, i] for i in range(ny)], dtype=int),
                "right": np.array([[nx - 1, i] for i in range(ny)], dtype=int)}

    # ---- hooks
    def set_boundary_conditions(self):
        pass

    def get_source(self):
        return None

    def initialize_macroscopic_fields(self):
        return None, None

    def output_data(self, **kwargs):
        pass

    # ---- darcy.py:36-73
    def update_macroscopic(self, f, external_source=None, gradient=False):
        rho = torch.sum(f, dim=-1, keepdim=True)
        if self.source is not None and external_source is not None:
            rho = rho + 0.5 * (self.source + external_source)
        elif self.source is not None:
            rho = rho + 0.5 * self.source
        elif external_source is not None:
            rho = rho + 0.5 * external_source
        if gradient:
            c = torch.as_tensor(self.c, dtype=self.precisionPolicy.compute_dtype).T
            om = self._omega(self.precisionPolicy.compute_dtype)
            if isinstance(om, float):
                grad = torch.matmul(f.to(c.dtype), (-3 * om) * c)
            else:
                grad = torch.matmul(f.to(c.dtype), c) * (-3 * om)
            return rho, grad
        return rho

    # ---- darcy.py:75-104
    def gradient(self, q, scale=1.0):
        p_in = q.repeat(1, 1, self.q)
        p_out = stream(p_in, self.lattice)
        for bc in self.BCs:
            if bc.name in ("ADR_Neumann_BounceBack", "ADR_AntiBounceBack"):
                ix, iy, keep = self._inbounds(bc, q.shape)
                vals = bc.apply(p_out, p_in) if keep is None else self._apply_clamped(bc, p_out, p_in, keep)
                p_out = p_out.clone()
                p_out[ix, iy] = vals.to(p_out.dtype)
        w = torch.as_tensor(self.w, dtype=q.dtype)
        c = torch.as_tensor(self.c, dtype=q.dtype)
        return torch.matmul(p_out, -3 * scale * w[..., None] * c.T)

    def _inbounds(self, bc, shape):
        """JAX ``.at[]`` semantics when fine-level BC indices meet a coarse array:
        out-of-range scatter updates are dropped (documented default)."""
        ix = np.asarray(bc.indices[0])
        iy = np.asarray(bc.indices[1])
        if ix.max(initial=0) < shape[0] and iy.max(initial=0) < shape[1]:
            return torch.as_tensor(ix), torch.as_tensor(iy), None
        keep = (ix < shape[0]) & (iy < shape[1])
        return torch.as_tensor(ix[keep]), torch.as_tensor(iy[keep]), keep

    def _apply_clamped(self, bc, p_out, p_in, keep):
        """Gathers clamp out-of-range indices (JAX default); only the kept rows survive."""
        saved = bc.indices
        ix = np.minimum(np.asarray(saved[0]), p_out.shape[0] - 1)
        iy = np.minimum(np.asarray(saved[1]), p_out.shape[1] - 1)
        bc.indices = (ix, iy)
        try:
            vals = bc.apply(p_out, p_in)
        finally:
            bc.indices = saved
        return vals[torch.as_tensor(keep)]

    # ---- darcy.py:106-129
    def collision(self, f, external_source=None):
        pp = self.precisionPolicy
        f = pp.cast_to_compute(f)
        rho = self.update_macroscopic(f, external_source)
        w = torch.as_tensor(self.w, dtype=f.dtype)
        feq = rho * w
        fneq = f - feq
        om = self._omega(f.dtype)
        fout = f - om * fneq
        if self.source is not None and external_source is not None:
            fout = fout + (self.source + external_source) * ((1 - 0.5 * om) * w)
        elif self.source is not None:
            fout = fout + self.source * ((1 - 0.5 * om) * w)
        elif external_source is not None:
            fout = fout + external_source * ((1 - 0.5 * om) * w)
        return pp.cast_to_output(fout)

    # ---- darcy.py:131-171
    def step(self, f_poststreaming, timestep, return_fpost=False, external_source=None):
        f_pc = self.collision(f_poststreaming, external_source)
        f_pc = obc.apply_bc(self.BCs, f_pc, f_poststreaming, "PostCollision")
        f_ps = stream(f_pc, self.lattice)
        f_ps = obc.apply_bc(self.BCs, f_ps, f_pc, "PostStreaming")
        return (f_ps, f_pc) if return_fpost else (f_ps, None)

    # ---- base.py:528-579, 685-724 (u = 0 equilibrium: rho * w)
    def assign_fields_sharded(self):
        rho0, u0 = self.initialize_macroscopic_fields()
        w = torch.as_tensor(self.w, dtype=self.precisionPolicy.compute_dtype)
        if rho0 is None or u0 is None:
            return w.expand(self.nx, self.ny, self.q).to(self.precisionPolicy.output_dtype).clone()
        rho0 = torch.as_tensor(np.asarray(rho0)).to(self.precisionPolicy.compute_dtype)
        u0 = torch.as_tensor(np.asarray(u0)).to(self.precisionPolicy.compute_dtype)
        c = torch.as_tensor(self.c, dtype=rho0.dtype)
        cu = 3.0 * torch.matmul(u0, c)
        usqr = 1.5 * torch.sum(u0 ** 2, dim=-1, keepdim=True)
        feq = rho0 * w * (1.0 + cu * (1.0 + 0.5 * cu) - usqr)
        return feq.to(self.precisionPolicy.output_dtype)

    # ---- darcy.py:173-252 (t_max steps, from timestep 1)
    def run(self, t_max):
        f = self.assign_fields_sharded()
        if self.ioRate > 0:
            self.output_data(timestep=0, rho=gathered(self.update_macroscopic(f)), rho_prev=None,
                             f_poststreaming=f, f_postcollision=None)
        for timestep in range(1, t_max + 1):
            io_flag = self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max)
            if io_flag:
                rho_prev = gathered(self.update_macroscopic(f))
            f, fstar = self.step(f, timestep, return_fpost=self.returnFpost)
            if io_flag:
                self.output_data(timestep=timestep, rho=gathered(self.update_macroscopic(f)), rho_prev=rho_prev,
                                 f_poststreaming=f, f_postcollision=fstar)
        return f
