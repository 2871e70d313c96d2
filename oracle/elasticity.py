"""D2Q8 moment-based LBM for static linear elasticity, with the multigrid V-cycle.

Oracle / test use only.  Restates ``src/elasticity.py`` of the reference in
torch-CPU ops on arrays of layout ``(nx, ny, q)``.  Subclasses override the same
hooks as the reference (``set_boundary_conditions``, ``get_force``,
``initialize_macroscopic_fields``, ``output_data``).
"""
import numpy as np
import torch

from . import bc as obc
from .bc import PrecisionPolicy, gathered, stream


class OracleElasticity:
    def __init__(self, **kwargs):
        # elasticity.py:51-148 (single device: no x padding, identity halo exchange)
        self.lamda = kwargs.get("lambda")
        self.mu = kwargs.get("mu")
        self.nx = kwargs.get("nx")
        self.ny = kwargs.get("ny")
        self.nz = kwargs.get("nz", 0)
        mg = {"left": False, "right": False, "bottom": False, "top": False, "front": False, "back": False}
        mg.update(kwargs.get("MG_BCs", {}))
        self.mg_BCs = mg
        self.levels = kwargs.get("MG_levels", 1)
        self.gamma = kwargs.get("MG_relaxation", 0.75)
        self.k_smooth = kwargs.get("MG_n_smoothing", 2)
        self.precision = kwargs.get("precision", "f64/f64")
        self.precisionPolicy = PrecisionPolicy(self.precision)
        self.lattice = kwargs.get("lattice")
        self.returnFpost = kwargs.get("return_fpost", False)
        self.ioRate = kwargs.get("io_rate", 0)
        self.c = self.lattice.c
        self.q = self.lattice.q
        self.w = self.lattice.w
        self.dim = 2
        den = 2 ** (self.levels - 1)
        if (self.nx - int(mg["left"]) - int(mg["right"])) % den or (self.ny - int(mg["bottom"]) - int(mg["top"])) % den:
            raise ValueError("nx,ny,nz (-BCs) must be divisible by 2**(levels-1)")
        self._create_transfer_matrices()
        self.gridInfo = self.grid_info(0)
        self.boundingBoxIndices = self.bounding_box_indices()
        self._create_boundary_data()
        force = self.get_force()
        self.force = None if force is None else torch.as_tensor(np.asarray(force), dtype=self.precisionPolicy.compute_dtype)
        # per-cell Lame parameters (nx, ny) (an extension of the reference's scalars, for
        # the gradient checks: tensors that may require grad)
        self.lam_field = kwargs.get("lambda_field")
        self.mu_field = kwargs.get("mu_field")

    def _fine(self, f):
        """Per-cell Lame fields apply to arrays of the finest grid; coarse multigrid
        levels use the scalar lambda, mu."""
        return self.lam_field is not None and tuple(f.shape[:2]) == (self.nx, self.ny)

    def _coef(self, fields=True):
        """(cs, cd, h, n) of the collision, scalars or per-cell (nx, ny, 1) tensors,
        from elasticity.py:811-813 and :976-979."""
        if self.lam_field is None or not fields:
            lm, mu = self.lamda + self.mu, self.mu
            g = 1.0 / (12.0 * lm - 4.0)
            g = 0.0 if abs(g) > 100 else g
        else:
            lm = (self.lam_field + self.mu_field)[..., None]
            mu = self.mu_field[..., None]
            g = 1.0 / (12.0 * lm - 4.0)
            g = torch.where(g.abs() > 100, torch.zeros_like(g), g)
        return (3 * lm - 1) / (3 * lm + 1), (6 * mu - 1) / (6 * mu + 1), 0.25 + 0.5 * g, -0.25 * g

    def _to_f(self, m, h, n):
        # m @ m_to_f with per-cell h, n in the column of m_s
        if not torch.is_tensor(h):
            return torch.matmul(m, self.m_to_f)
        base = self.m_to_f.clone()
        base[2, :] = 0.0
        hn = torch.cat([h.expand(*m.shape[:-1], 4), n.expand(*m.shape[:-1], 4)], dim=-1)
        return torch.matmul(m, base) + m[..., 2:3] * hn

    # ---- geometry (elasticity.py:501-604)
    def grid_size(self, level=0):
        if level == 0:
            return self.nx, self.ny
        nxb = int(self.mg_BCs["left"]) + int(self.mg_BCs["right"])
        nyb = int(self.mg_BCs["bottom"]) + int(self.mg_BCs["top"])
        return (self.nx - nxb) // 2 ** level + nxb, (self.ny - nyb) // 2 ** level + nyb

    def grid_info(self, level=0):
        nx, ny = self.grid_size(level)
        return {"nx": nx, "ny": ny, "nz": 0, "dim": 2, "lattice": self.lattice}

    def bounding_box_indices(self, level=0):
        nx, ny = self.grid_size(level)
        return {"bottom": np.array([[i, 0] for i in range(nx)], dtype=int),
                "top": np.array([[i, ny - 1] for i in range(nx)], dtype=int),
                "left": np.array([[0, i] for i in range(ny)], dtype=int),
                "right": np.array([[nx - 1, i] for i in range(ny)], dtype=int)}

    def _create_boundary_data(self):
        # elasticity.py:396-436
        self.BCs = []
        self.BCs_MG = {level: [] for level in range(1, self.levels)}
        self.set_boundary_conditions()
        obc.build_bcs(self.BCs, self.nx, self.ny, self.lattice)
        for level in range(1, self.levels):
            nx, ny = self.grid_size(level)
            obc.build_bcs(self.BCs_MG[level], nx, ny, self.lattice)

    # ---- hooks (elasticity.py:631-645, 1458-1501)
    def set_boundary_conditions(self):
        pass

    def get_force(self):
        return None

    def initialize_macroscopic_fields(self):
        return None, None

    def output_data(self, **kwargs):
        pass

    # ---- constants (elasticity.py:805-848)
    def _create_transfer_matrices(self):
        g = 1.0 / (12.0 * (self.lamda + self.mu) - 4.0)
        if abs(g) > 100:
            g = 0.0
        dt = self.precisionPolicy.compute_dtype
        self.f_to_m = torch.tensor([[1, 0, -1, 0, 1, -1, -1, 1],
                                    [0, 1, 0, -1, 1, 1, -1, -1],
                                    [1, 1, 1, 1, 2, 2, 2, 2],
                                    [1, -1, 1, -1, 0, 0, 0, 0],
                                    [0, 0, 0, 0, 1, -1, 1, -1]], dtype=dt).T
        h = 0.25 + 0.5 * g
        n = -0.25 * g
        self.m_to_f = torch.tensor([[0.5, 0, h, 0.25, 0, -0.5, 0],
                                    [0, 0.5, h, -0.25, 0, 0, -0.5],
                                    [-0.5, 0, h, 0.25, 0, 0.5, 0],
                                    [0, -0.5, h, -0.25, 0, 0, 0.5],
                                    [0, 0, n, 0, 0.25, 0.25, 0.25],
                                    [0, 0, n, 0, -0.25, -0.25, 0.25],
                                    [0, 0, n, 0, 0.25, -0.25, -0.25],
                                    [0, 0, n, 0, -0.25, 0.25, -0.25]], dtype=dt).T

    # ---- pointwise maps
    def equilibrium(self, eta, sigma, external_force=None):
        # elasticity.py:908-957
        pp = self.precisionPolicy
        eta = torch.as_tensor(eta).to(pp.compute_dtype)
        sigma = torch.as_tensor(sigma).to(pp.compute_dtype)
        if self.force is not None:
            eta = eta - 0.5 * self.force
        if external_force is not None:
            eta = eta - 0.5 * torch.as_tensor(external_force).to(pp.compute_dtype)
        if self.lam_field is None:
            lm, mu = self.lamda + self.mu, self.mu
        else:
            lm, mu = (self.lam_field + self.mu_field)[..., None], self.mu_field[..., None]
        cs = -1.0 - 1 / (3 * lm)
        ms = (sigma[..., 0:1] + sigma[..., 2:3]) * cs
        cd = -1.0 - 1 / (6 * mu)
        md = (sigma[..., 0:1] - sigma[..., 2:3]) * cd
        m11 = sigma[..., 1:2] * cd
        m = torch.cat((eta, ms, md, m11, eta / 3.0), dim=-1)
        _, _, h, n = self._coef()
        return pp.cast_to_output(self._to_f(m, h, n))

    def collision(self, f, external_force, use_force=True):
        # elasticity.py:959-1000
        pp = self.precisionPolicy
        f = pp.cast_to_compute(f)
        external_force = pp.cast_to_compute(external_force)
        m = torch.matmul(f, self.f_to_m)
        cs, cd, h, n = self._coef(self._fine(f))
        if use_force and self.force is not None and external_force is not None:
            force = self.force + external_force
        elif external_force is not None:
            force = external_force
        elif use_force and self.force is not None:
            force = self.force
        else:
            force = None
        if force is not None:
            m = torch.cat([m[..., 0:2] + force, m[..., 2:3] * cs, m[..., 3:5] * cd,
                           m[..., 0:2] / 3.0 + force / 6.0], dim=-1)
        else:
            m = torch.cat([m[..., 0:2], m[..., 2:3] * cs, m[..., 3:5] * cd, m[..., 0:2] / 3.0], dim=-1)
        return pp.cast_to_output(self._to_f(m, h, n))

    def update_macroscopic(self, f, external_force=None):
        # elasticity.py:1002-1038
        m = torch.matmul(f.to(self.f_to_m.dtype), self.f_to_m)
        eta = m[..., 0:2]
        if self.force is not None:
            eta = eta + 0.5 * self.force
        if external_force is not None:
            eta = eta + 0.5 * external_force
        if self.lam_field is None:
            lm, mu = self.lamda + self.mu, self.mu
        else:
            lm, mu = self.lam_field + self.mu_field, self.mu_field
        cs = -1.5 * lm / (3 * lm + 1)
        ms = m[..., 2] * cs
        cd = -3 * mu / (6 * mu + 1)
        md = m[..., 3] * cd
        sigma = torch.stack((ms + md, (2 * cd) * m[..., 4], ms - md), dim=-1)
        return eta, sigma

    def update_divergence(self, f):
        # elasticity.py:1040-1058
        lm = (self.lam_field + self.mu_field)[..., None] if self._fine(f) else self.lamda + self.mu
        return (-1.5 / (3 * lm + 1)) * torch.matmul(f.to(self.f_to_m.dtype), self.f_to_m[:, 2:3])

    # ---- step (elasticity.py:1128-1219)
    def _bcs(self, level):
        return self.BCs if level == 0 else self.BCs_MG[level]

    def step(self, f_poststreaming, timestep, return_fpost=False, external_force=None, nonlinear_force=None):
        if self.levels == 1:
            if callable(nonlinear_force):
                external_force = external_force + nonlinear_force(f_poststreaming)
            return self._step(f_poststreaming, timestep, return_fpost, external_force)
        return self._mg_step(f_poststreaming, timestep, external_force=external_force, nonlinear_force=nonlinear_force)

    def _step(self, f_poststreaming, timestep, return_fpost=False, external_force=None, level=0):
        f_pc = self.collision(f_poststreaming, external_force, level == 0)
        f_pc = obc.apply_bc(self._bcs(level), f_pc, f_poststreaming, "PostCollision")
        f_ps = stream(f_pc, self.lattice)
        f_ps = obc.apply_bc(self._bcs(level), f_ps, f_pc, "PostStreaming")
        return (f_ps, f_pc) if return_fpost else (f_ps, None)

    # ---- multigrid (elasticity.py:850-906, 1221-1344)
    def _box(self, shape):
        return (int(self.mg_BCs["left"]), shape[0] - int(self.mg_BCs["right"]),
                int(self.mg_BCs["bottom"]), shape[1] - int(self.mg_BCs["top"]))

    def _restrict(self, f):
        x0, x1, y0, y1 = self._box(f.shape)
        a = f[x0:x1, y0:y1]
        return a[0::2, 0::2] + a[1::2, 0::2] + a[0::2, 1::2] + a[1::2, 1::2]

    def _prolong(self, f):
        x0, x1, y0, y1 = self._box(f.shape)
        a = f[x0:x1, y0:y1]
        return a.repeat_interleave(2, dim=0).repeat_interleave(2, dim=1)

    def substract_residual(self, f, residual):
        if residual is None:
            return f
        x0, x1, y0, y1 = self._box(f.shape)
        f = f.clone()
        f[x0:x1, y0:y1] = f[x0:x1, y0:y1] + (-residual)
        return f

    def _ext(self, external_force, nonlinear_force, f):
        if callable(nonlinear_force):
            nl = nonlinear_force(f)
            return nl if external_force is None else external_force + nl
        return external_force

    def _mg_step(self, f_poststreaming, timestep, level=0, residual=None, external_force=None, nonlinear_force=None):
        g = self.gamma
        for _ in range(self.k_smooth):
            ext = self._ext(external_force, nonlinear_force, f_poststreaming)
            f_poststreaming = (1 - g) * f_poststreaming + g * self._step(
                self.substract_residual(f_poststreaming, residual), timestep, False, ext, level)[0]
        if level < self.levels - 1:
            x0, x1, y0, y1 = self._box(f_poststreaming.shape)
            ext = self._ext(external_force, nonlinear_force, f_poststreaming)
            res = f_poststreaming - self._step(self.substract_residual(f_poststreaming.clone(), residual),
                                               timestep, False, ext, level)[0]
            res_c = self._restrict(res)
            f_c = torch.zeros((*self.grid_size(level + 1), self.q), dtype=self.precisionPolicy.compute_dtype)
            f_c = self._mg_step(f_c, timestep, level + 1, res_c, nonlinear_force=nonlinear_force)
            f_poststreaming = f_poststreaming.clone()
            f_poststreaming[x0:x1, y0:y1] = f_poststreaming[x0:x1, y0:y1] + self._prolong(f_c)
        for _ in range(self.k_smooth):
            ext = self._ext(external_force, nonlinear_force, f_poststreaming)
            f_poststreaming = (1 - g) * f_poststreaming + g * self._step(
                self.substract_residual(f_poststreaming, residual), timestep, False, ext, level)[0]
        if level == 0:
            return f_poststreaming, torch.linalg.norm(res[x0:x1, y0:y1].reshape(-1))
        return f_poststreaming

    # ---- driver (elasticity.py:647-679, 1346-1425)
    def assign_fields_sharded(self):
        eta0, sigma0 = self.initialize_macroscopic_fields()
        if eta0 is None or sigma0 is None:
            return torch.zeros((self.nx, self.ny, self.q), dtype=self.precisionPolicy.output_dtype)
        return self.equilibrium(eta0, sigma0)

    def run(self, t_max):
        f = self.assign_fields_sharded()
        for timestep in range(0, t_max + 1):
            io_flag = self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max)
            if io_flag:
                eta_prev, sigma_prev = (gathered(x) for x in self.update_macroscopic(f))
            f, fstar = self.step(f, timestep, return_fpost=self.returnFpost)
            if io_flag:
                eta, sigma = (gathered(x) for x in self.update_macroscopic(f))
                self.output_data(timestep=timestep, eta=eta, eta_prev=eta_prev, sigma=sigma,
                                 sigma_prev=sigma_prev, f_poststreaming=f, f_postcollision=fstar)
        return f
