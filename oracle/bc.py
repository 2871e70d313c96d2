"""Boundary-condition machinery of the reference, restated on the CPU.  Oracle only.

Follows ``src/boundary_conditions.py:5-233`` (base class), ``:452-547``
(``BounceBackHalfway``), ``src/elasticity.py:1503-1750`` (elastic BCs) and
``src/darcy.py:298-466`` (ADR BCs), plus the grid mask of
``src/elasticity.py:460-499`` / ``src/base.py:413-449``.

Scatter semantics.  The reference writes ``fbd.at[bindex, imissing].set(v)`` where
``imissing`` holds 0 for every non-missing direction, so several updates hit
column 0 of a row.  XLA:CPU applies scatter updates in row-major order, so the
last update wins for ``set`` and all updates accumulate for ``add``.  Rows are
distinct cells (indices are ``np.unique``-d in ``configure``), hence updates of
different rows never collide and the row-major order reduces to "column k
ascending", which is what :func:`scatter_set` / :func:`scatter_add` do.
"""
import numpy as np
import torch


class PrecisionPolicy:
    """Stand-in for ``jmp.Policy`` (``elasticity.py:63-66``, ``:606-629``)."""
    _TABLE = {
        "f64/f64": (torch.float64, torch.float64),
        "f32/f32": (torch.float32, torch.float32),
        "f64/f32": (torch.float64, torch.float32),
        "f64/f16": (torch.float64, torch.float16),
        "f32/f16": (torch.float32, torch.float16),
    }

    def __init__(self, precision):
        self.compute_dtype, self.output_dtype = self._TABLE.get(precision, (torch.float32, torch.float32))

    def cast_to_compute(self, x):
        return None if x is None else x.to(self.compute_dtype)

    def cast_to_output(self, x):
        return None if x is None else x.to(self.output_dtype)


def gathered(x):
    """A field as the run loops hand it to ``output_data``: on the host with the leading
    process axis of ``process_allgather`` (elasticity.py:1378-1401, darcy.py:200-229,
    poroelasticity.py:273-277; one process)."""
    return x.detach().cpu().numpy()[None]


def stream(f, lattice):
    """``streaming_p`` (elasticity.py:768-803): ``f_i(x) <- f_i(x - c_i)`` periodic roll.

    The single-device ``ppermute`` fix-up of ``streaming_m`` (:738-766) is the identity."""
    out = torch.empty_like(f)
    for i in range(lattice.q):
        out[..., i] = torch.roll(f[..., i], shifts=(int(lattice.c[0, i]), int(lattice.c[1, i])), dims=(0, 1))
    return out


def create_grid_mask(nx, ny, lattice, solid_voxels):
    """elasticity.py:460-487 with one device: halo width 1, halo + solid voxels True, streamed once."""
    mask = np.ones((nx + 2, ny + 2, lattice.q), dtype=bool)
    mask[1:-1, 1:-1, :] = False
    if solid_voxels is not None:
        sv = np.array(solid_voxels)
        mask[sv[:, 0] + 1, sv[:, 1] + 1, :] = True
    out = np.empty_like(mask)
    for i in range(lattice.q):
        out[..., i] = np.roll(mask[..., i], (int(lattice.c[0, i]), int(lattice.c[1, i])), axis=(0, 1))
    return out


def scatter_set(fbd, cols, vals):
    """``fbd.at[bindex, cols].set(vals)`` with XLA:CPU in-order semantics (last write wins)."""
    rows = torch.arange(fbd.shape[0])
    for k in range(cols.shape[1]):
        fbd[rows, torch.as_tensor(cols[:, k], dtype=torch.long)] = vals[:, k]
    return fbd


def scatter_add(fbd, cols, vals):
    """``fbd.at[bindex, cols].add(vals)``: every update accumulates."""
    rows = torch.arange(fbd.shape[0])
    for k in range(cols.shape[1]):
        idx = torch.as_tensor(cols[:, k], dtype=torch.long)
        fbd[rows, idx] = fbd[rows, idx] + vals[:, k]
    return fbd


def gather_cols(a, cols):
    return torch.gather(a, 1, torch.as_tensor(np.asarray(cols), dtype=torch.long))


class BoundaryCondition:
    """boundary_conditions.py:41-233."""

    def __init__(self, indices, gridInfo, precision_policy):
        self.lattice = gridInfo["lattice"]
        self.nx = gridInfo["nx"]
        self.ny = gridInfo["ny"]
        self.nz = gridInfo["nz"]
        self.dim = gridInfo["dim"]
        self.precisionPolicy = precision_policy
        self.indices = tuple(np.asarray(a) for a in indices)
        self.name = None
        self.isSolid = False
        self.isDynamic = False
        self.needsExtraConfiguration = False
        self.implementationStep = "PostStreaming"

    def create_local_mask_and_normal_arrays(self, grid_mask):
        # boundary_conditions.py:55-85
        if self.needsExtraConfiguration:
            boundaryMask = self.get_boundary_mask(grid_mask)
            self.configure(boundaryMask)
            self.needsExtraConfiguration = False
        boundaryMask = self.get_boundary_mask(grid_mask)
        self.normals = self.get_normals(boundaryMask)
        self.imissing, self.iknown = self.get_missing_indices(boundaryMask)
        self.imissingMask, self.iknownMask, self.imiddleMask = self.get_missing_mask(boundaryMask)

    def get_boundary_mask(self, grid_mask):
        # :87-109 (device_count() == 1)
        sx = np.asarray(self.indices[0]) + 1
        sy = np.asarray(self.indices[1]) + 1
        return np.array(grid_mask[sx, sy])

    def configure(self, boundaryMask):
        return

    def get_normals(self, boundaryMask):
        # :155-176
        main_c = self.lattice.c.T[self.lattice.main_indices]
        m = boundaryMask[..., self.lattice.main_indices]
        return -np.dot(m, main_c)

    def get_missing_indices(self, boundaryMask):
        # :178-206
        nbd = len(self.indices[0])
        imissing = np.vstack([np.arange(self.lattice.q, dtype='uint8')] * nbd)
        iknown = np.vstack([self.lattice.opp_indices] * nbd)
        imissing[~boundaryMask] = 0
        iknown[~boundaryMask] = 0
        return imissing, iknown

    def get_missing_mask(self, boundaryMask):
        # :208-233
        imissingMask = boundaryMask
        iknownMask = imissingMask[:, self.lattice.opp_indices]
        imiddleMask = ~(imissingMask | iknownMask)
        return imissingMask, iknownMask, imiddleMask

    def _shift_to_fluid(self, boundaryMask):
        """Index shift shared by every halfway BC ``configure``
        (elasticity.py:1595-1603, darcy.py:338-346, boundary_conditions.py:499-507)."""
        hasFluidNeighbour = ~boundaryMask[:, self.lattice.opp_indices]
        idx = np.array(self.indices).T
        idx_trg = []
        for i in range(self.lattice.q):
            idx_trg.append(idx[hasFluidNeighbour[:, i], :] + self.lattice.c[:, i])
        indices_new, array_indices = np.unique(np.vstack(idx_trg), axis=0, return_index=True)
        return indices_new, array_indices

    def _rows(self, a):
        return a[torch.as_tensor(self.indices[0], dtype=torch.long), torch.as_tensor(self.indices[1], dtype=torch.long)]


def _cu(lattice, vec, dtype):
    """``6 w_i (c_i . v)`` (elasticity.py:1548-1549, :1616-1617; boundary_conditions.py:518-519)."""
    c = torch.as_tensor(lattice.c, dtype=dtype)
    w = torch.as_tensor(lattice.w, dtype=dtype)
    return 6.0 * w * torch.matmul(torch.as_tensor(vec, dtype=dtype), c)


class ElasticityBounceBackFullway(BoundaryCondition):
    """elasticity.py:1503-1550: PostCollision, ``f*_i = f_opp(i)(pre-collision) + 6 w_i c_i.eta``."""

    def __init__(self, indices, gridInfo, precision_policy, eta=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityBounceBackFullway"
        self.implementationStep = "PostCollision"
        self.eta = eta

    def apply(self, fout, fin):
        base = self._rows(fin)[:, torch.as_tensor(self.lattice.opp_indices)]
        if self.eta is None:
            return base
        return base + _cu(self.lattice, self.eta, fin.dtype)


class ElasticityBounceBack(BoundaryCondition):
    """elasticity.py:1552-1646: halfway Dirichlet, PostStreaming."""

    def __init__(self, indices, gridInfo, precision_policy, eta=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.eta = None if eta is None else np.asarray(eta, dtype=np.float64)

    def configure(self, boundaryMask):
        # :1581-1612
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.eta is not None:
            if self.eta.shape == (self.lattice.d,):
                self.eta = np.tile(self.eta, (nbd_modified, 1))
            elif nbd_orig != nbd_modified:
                self.eta = self.eta[array_indices]

    def apply(self, fout, fin):
        # :1622-1646 and impose_boundary_displacement :1614-1620
        fbd = self._rows(fout).clone()
        fin_bd = self._rows(fin)
        scatter_set(fbd, self.imissing, gather_cols(fin_bd, self.iknown))
        if self.eta is not None:
            cu = _cu(self.lattice, self.eta, fout.dtype)
            scatter_add(fbd, self.imissing, gather_cols(cu, self.imissing))
        return fbd


class ElasticityAntiBounceBack(BoundaryCondition):
    """elasticity.py:1648-1750: halfway Neumann (traction), PostStreaming.

    ``fbd.at[imissingMask].set(-fin[iknownMask])`` pairs the k-th True of the missing
    mask with the k-th True of the known mask in row-major order."""

    def __init__(self, indices, gridInfo, precision_policy, T=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ElasticityAntiBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.T = None if T is None else np.asarray(T, dtype=np.float64)

    def configure(self, boundaryMask):
        # :1677-1708
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.T is not None:
            if self.T.shape == (self.lattice.d,):
                self.T = np.tile(self.T, (nbd_modified, 1))
            elif nbd_orig != nbd_modified:
                self.T = self.T[array_indices]

    def get_missing_mask(self, boundaryMask):
        # :1710-1727
        if self.T is not None:
            normals = self.get_normals(boundaryMask)
            T = self.T
            if np.all(normals[:, 0] == -1):
                self.i_update = (0, 4, 7)
                self.T_update = np.stack((T[:, 0], 0.5 * T[:, 1], -0.5 * T[:, 1]), axis=-1)
            elif np.all(normals[:, 0] == 1):
                self.i_update = (2, 5, 6)
                self.T_update = np.stack((-T[:, 0], 0.5 * T[:, 1], -0.5 * T[:, 1]), axis=-1)
            elif np.all(normals[:, 1] == -1):
                self.i_update = (1, 4, 5)
                self.T_update = np.stack((T[:, 1], 0.5 * T[:, 0], -0.5 * T[:, 0]), axis=-1)
            elif np.all(normals[:, 1] == 1):
                self.i_update = (3, 6, 7)
                self.T_update = np.stack((-T[:, 1], -0.5 * T[:, 0], 0.5 * T[:, 0]), axis=-1)
            else:
                raise NotImplementedError
        return super().get_missing_mask(boundaryMask)

    def apply(self, fout, fin):
        # :1730-1750
        fbd = self._rows(fout).clone()
        fin_bd = self._rows(fin)
        mm = torch.as_tensor(self.imissingMask)
        km = torch.as_tensor(self.iknownMask)
        fbd[mm] = -fin_bd[km]
        if self.T is not None:
            cols = torch.as_tensor(self.i_update)
            fbd[:, cols] = fbd[:, cols] + torch.as_tensor(self.T_update, dtype=fbd.dtype)
        return fbd


class BounceBackHalfway(BoundaryCondition):
    """boundary_conditions.py:452-547 (``vel=None`` on this path)."""

    def __init__(self, indices, gridInfo, precision_policy, vel=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "BounceBackHalfway"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.vel = vel

    def configure(self, boundaryMask):
        indices_new, _ = self._shift_to_fluid(boundaryMask)
        nbd_orig = len(self.indices[0])
        self.indices = tuple(indices_new.T)
        if self.vel is not None and nbd_orig != len(self.indices[0]):
            vel_avg = np.mean(np.asarray(self.vel), axis=0)
            self.vel = np.zeros(indices_new.shape) + vel_avg

    def apply(self, fout, fin):
        fbd = self._rows(fout).clone()
        fin_bd = self._rows(fin)
        scatter_set(fbd, self.imissing, gather_cols(fin_bd, self.iknown))
        if self.vel is not None:
            cu = _cu(self.lattice, self.vel, fout.dtype)
            scatter_add(fbd, self.imissing, -gather_cols(cu, self.iknown))
        return fbd


class ADR_AntiBounceBack(BoundaryCondition):
    """darcy.py:298-392: halfway anti-bounce-back (Dirichlet ``rho``), PostStreaming."""

    def __init__(self, indices, gridInfo, precision_policy, rho=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ADR_AntiBounceBack"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.rho = rho

    def configure(self, boundaryMask):
        # :326-360
        nbd_orig = len(self.indices[0])
        indices_new, array_indices = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        nbd_modified = len(self.indices[0])
        if self.rho is not None:
            if isinstance(self.rho, float) or np.asarray(self.rho).size == 1:
                # the reference only broadcasts when the count changed; a scalar that
                # survives would fail at apply time there, here it is broadcast as well
                self.rho = np.full((nbd_modified, 1), float(np.asarray(self.rho).ravel()[0]))
            elif nbd_orig != nbd_modified:
                self.rho = np.asarray(self.rho)[array_indices]

    def apply(self, fout, fin):
        # :362-392
        fbd = self._rows(fout).clone()
        fin_bd = self._rows(fin)
        scatter_set(fbd, self.imissing, -gather_cols(fin_bd, self.iknown))
        if self.rho is not None:
            tmp = 2 * torch.as_tensor(self.lattice.w, dtype=fbd.dtype) * torch.as_tensor(
                np.asarray(self.rho, dtype=np.float64).reshape(-1, 1), dtype=fbd.dtype)
            scatter_add(fbd, self.imissing, gather_cols(tmp, self.iknown))
        return fbd


class ADR_Neumann_BounceBack(BounceBackHalfway):
    """darcy.py:394-415."""

    def __init__(self, indices, gridInfo, precision_policy):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "ADR_Neumann_BounceBack"


class ADR_EquilibriumBC(BoundaryCondition):
    """darcy.py:417-466 (``vel=None``): BC cells set to ``rho * w`` after streaming."""

    def __init__(self, indices, gridInfo, precision_policy, rho, vel=None):
        super().__init__(indices, gridInfo, precision_policy)
        rho = torch.as_tensor(np.asarray(rho, dtype=np.float64))
        w = torch.as_tensor(self.lattice.w, dtype=torch.float64)
        self.out = precision_policy.cast_to_output(rho.reshape(-1, 1) * w if rho.dim() > 0 else rho * w)
        self.name = "ADR_EquilibriumBC"
        self.implementationStep = "PostStreaming"

    def apply(self, fout, fin):
        nbd = len(self.indices[0])
        return self.out.to(fout.dtype).expand(nbd, self.lattice.q)


def apply_bc(bcs, fout, fin, implementation_step):
    """elasticity.py:1060-1091 / base.py:775-806: BCs in append order, last write wins."""
    for bc in bcs:
        if bc.implementationStep == implementation_step:
            ix = torch.as_tensor(bc.indices[0], dtype=torch.long)
            iy = torch.as_tensor(bc.indices[1], dtype=torch.long)
            fout = fout.clone()
            fout[ix, iy] = bc.apply(fout, fin).to(fout.dtype)
    return fout


def build_bcs(bcs, nx, ny, lattice):
    """``_create_boundary_data`` (elasticity.py:396-436, base.py:351-372)."""
    solid = [np.array(bc.indices).T for bc in bcs if bc.isSolid]
    solid_voxels = np.unique(np.vstack(solid), axis=0) if solid else None
    grid_mask = create_grid_mask(nx, ny, lattice, solid_voxels)
    for bc in bcs:
        assert bc.implementationStep in ("PostStreaming", "PostCollision")
        bc.create_local_mask_and_normal_arrays(grid_mask)
