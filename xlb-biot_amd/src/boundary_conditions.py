"""Boundary conditions: the reference's construction machinery plus a rule compiler.

The construction half keeps the reference API and call order of
``src/boundary_conditions.py:41-233`` (``create_local_mask_and_normal_arrays``
-> ``get_boundary_mask`` -> ``configure`` -> ``get_normals`` ->
``get_missing_indices`` -> ``get_missing_mask``), so user subclasses that
override one of these hooks (the drivers' corner fixes) behave as in the reference.

The run-time half is new.  Instead of gather/scatter passes per step, every BC
object symbolically applies its ``apply`` to a per-cell state of *rules*
(:meth:`BoundaryCondition.apply_rules`), in append order, exactly as
``fout.at[bc.indices].set(bc.apply(fout, fin))`` would (elasticity.py:1083-1089).
The result is one rule per (boundary cell, direction):

    value = base + const,   base in {default, +src[k], -src[k], 0}

which the fused kernels evaluate in registers.  The scatter semantics follow
XLA:CPU: ``.at[rows, cols].set`` applies updates in row-major order (last write
wins), ``.add`` accumulates every update.  This matters because the reference's
``imissing`` holds index 0 for every non-missing direction.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat

KEEP = 0
ZERO = nat.XB_RULE_ZERO


class RuleState:
    """Symbolic values of ``fbd`` rows: ``code`` (nbd, q) and ``const`` (nbd, q)."""

    def __init__(self, code, const):
        self.code = code
        self.const = const

    def set_from_fin(self, cols, src_cols, sign):
        """``fbd.at[bindex, cols].set(sign * fin[bindex, src_cols])`` (row-major, last wins)."""
        rows = np.arange(self.code.shape[0])
        cols = np.asarray(cols, dtype=np.int64)
        src_cols = np.asarray(src_cols, dtype=np.int64)
        for k in range(cols.shape[1]):
            self.code[rows, cols[:, k]] = sign * (src_cols[:, k] + 1)
            self.const[rows, cols[:, k]] = 0.0

    def add(self, cols, vals):
        """``fbd.at[bindex, cols].add(vals)``: every update accumulates."""
        rows = np.arange(self.code.shape[0])
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        for k in range(cols.shape[1]):
            np.add.at(self.const, (rows, cols[:, k]), vals[:, k])

    def set_masked_from_fin(self, missing_mask, known_mask, sign):
        """``fbd.at[missingMask].set(sign * fin[knownMask])``: k-th True <- k-th True, row-major."""
        mr, mc = np.nonzero(missing_mask)
        kr, kc = np.nonzero(known_mask)
        if len(mr) != len(kr):
            raise ValueError("missing and known masks select different counts")
        self.code[mr, mc] = sign * (kc + 1)
        self.const[mr, mc] = 0.0

    def set_const(self, vals):
        self.code[:, :] = ZERO
        self.const[:, :] = np.broadcast_to(np.asarray(vals, dtype=np.float64), self.const.shape)


class BoundaryCondition(object):
    """Base class (boundary_conditions.py:5-233)."""

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
        if self.needsExtraConfiguration:
            boundaryMask = self.get_boundary_mask(grid_mask)
            self.configure(boundaryMask)
            self.needsExtraConfiguration = False
        boundaryMask = self.get_boundary_mask(grid_mask)
        self.normals = self.get_normals(boundaryMask)
        self.imissing, self.iknown = self.get_missing_indices(boundaryMask)
        self.imissingMask, self.iknownMask, self.imiddleMask = self.get_missing_mask(boundaryMask)

    def get_boundary_mask(self, grid_mask):
        # one device: the x halo is one cell wide, like the y halo (boundary_conditions.py:102-107)
        return np.array(grid_mask[np.asarray(self.indices[0]) + 1, np.asarray(self.indices[1]) + 1])

    def configure(self, boundaryMask):
        return

    def prepare_populations(self, fout, fin, implementation_step):
        return fout

    def get_normals(self, boundaryMask):
        main = np.asarray(self.lattice.main_indices, dtype=np.int64)
        main_c = self.lattice.c.T.astype(np.int64)[main]
        return -np.dot(boundaryMask[..., main].astype(np.int64), main_c)

    def get_missing_indices(self, boundaryMask):
        nbd = len(self.indices[0])
        imissing = np.vstack([np.arange(self.lattice.q, dtype='uint8')] * nbd)
        iknown = np.vstack([np.asarray(self.lattice.opp_indices)] * nbd)
        imissing[~boundaryMask] = 0
        iknown[~boundaryMask] = 0
        return imissing, iknown

    def get_missing_mask(self, boundaryMask):
        imissingMask = boundaryMask
        iknownMask = imissingMask[:, np.asarray(self.lattice.opp_indices, dtype=np.int64)]
        imiddleMask = ~(imissingMask | iknownMask)
        return imissingMask, iknownMask, imiddleMask

    def _shift_to_fluid(self, boundaryMask):
        """Index shift to the fluid neighbours, shared by the halfway BCs
        (boundary_conditions.py:499-507, elasticity.py:1595-1603, darcy.py:338-346)."""
        opp = np.asarray(self.lattice.opp_indices, dtype=np.int64)
        hasFluidNeighbour = ~boundaryMask[:, opp]
        idx = np.array(self.indices).T
        c = self.lattice.c.astype(np.int64)
        idx_trg = [idx[hasFluidNeighbour[:, i], :] + c[:, i] for i in range(self.lattice.q)]
        return np.unique(np.vstack(idx_trg), axis=0, return_index=True)

    def apply_rules(self, state):
        """Symbolic ``apply``: update ``state`` (a :class:`RuleState` of this BC's rows)."""
        raise NotImplementedError(f"{type(self).__name__} has no rule form")

    def cu_rules(self, vec):
        """``6 w_i (c_i . v)`` per row (elasticity.py:1548-1549, boundary_conditions.py:518-519)."""
        c = self.lattice.c.astype(np.float64)
        w = np.asarray(self.lattice.w, dtype=np.float64)
        return 6.0 * w * np.dot(np.asarray(vec, dtype=np.float64), c)


class BounceBackHalfway(BoundaryCondition):
    """boundary_conditions.py:452-547."""

    def __init__(self, indices, gridInfo, precision_policy, vel=None):
        super().__init__(indices, gridInfo, precision_policy)
        self.name = "BounceBackHalfway"
        self.implementationStep = "PostStreaming"
        self.needsExtraConfiguration = True
        self.isSolid = True
        self.vel = vel

    def configure(self, boundaryMask):
        nbd_orig = len(self.indices[0])
        indices_new, _ = self._shift_to_fluid(boundaryMask)
        self.indices = tuple(indices_new.T)
        if self.vel is not None and nbd_orig != len(self.indices[0]):
            self.vel = np.zeros(indices_new.shape) + np.mean(np.asarray(self.vel), axis=0)

    def apply_rules(self, state):
        state.set_from_fin(self.imissing, self.iknown, +1)
        if self.vel is not None:
            cu = np.broadcast_to(self.cu_rules(self.vel), state.const.shape)
            state.add(self.imissing, -np.take_along_axis(cu, np.asarray(self.iknown, dtype=np.int64), 1))


# ---------------------------------------------------------------------------
# Host "BC compiler": grid mask, per-level configuration, rule tables.

def create_grid_mask(nx, ny, lattice, solid_voxels):
    """elasticity.py:460-487 / base.py:413-441 on one device: halo and solid voxels
    True, then streamed once (periodic roll of each direction plane)."""
    mask = np.ones((nx + 2, ny + 2, lattice.q), dtype=bool)
    mask[1:-1, 1:-1, :] = False
    if solid_voxels is not None:
        sv = np.asarray(solid_voxels)
        mask[sv[:, 0] + 1, sv[:, 1] + 1, :] = True
    out = np.empty_like(mask)
    c = lattice.c.astype(np.int64)
    for i in range(lattice.q):
        out[..., i] = np.roll(mask[..., i], (int(c[0, i]), int(c[1, i])), axis=(0, 1))
    return out


def configure_bcs(bcs, nx, ny, lattice):
    """``_create_boundary_data`` body (elasticity.py:405-419, base.py:356-370)."""
    solid = [np.array(bc.indices).T for bc in bcs if bc.isSolid]
    solid_voxels = np.unique(np.vstack(solid), axis=0) if solid else None
    grid_mask = create_grid_mask(nx, ny, lattice, solid_voxels)
    for bc in bcs:
        if bc.implementationStep not in ("PostStreaming", "PostCollision"):
            raise ValueError(f"unsupported implementation step {bc.implementationStep!r}")
        if bc.isDynamic:
            raise NotImplementedError("dynamic boundary conditions are not on this path")
        bc.create_local_mask_and_normal_arrays(grid_mask)


class RuleTable:
    """Device rule table of one (level, BC set, shape); owns the tensors behind the
    ``xb_bc_table`` struct handed to the kernels."""

    def __init__(self, nx, ny, q, slots, code_pc, const_pc, code_ps, const_ps, dtype, device):
        if torch.cuda.is_available() and torch.cuda.is_current_stream_capturing():
            # the tables are copied from host memory: build them before any capture
            raise RuntimeError("RuleTable must be created before a graph capture, not inside one")
        self.nx, self.ny, self.q = nx, ny, q
        self.ncells = int(len(slots))
        # cells whose rule rows (codes and constants, compared bit for bit) coincide
        # share one slot: walls with uniform data need only a handful of slots, which
        # the kernels then keep in LDS
        cols = []
        for code, const in ((code_pc, const_pc), (code_ps, const_ps)):
            if code is not None:
                cols.append(np.asarray(code, dtype=np.int64))
                cols.append(np.ascontiguousarray(np.asarray(const, dtype=np.float64)).view(np.int64))
        if self.ncells and cols:
            _, first, inverse = np.unique(np.concatenate(cols, axis=1), axis=0, return_index=True,
                                          return_inverse=True)
            inverse = inverse.reshape(-1)
        else:
            first, inverse = np.arange(self.ncells), np.arange(self.ncells)
        pick = lambda a: None if a is None else np.asarray(a)[first]
        code_pc, const_pc, code_ps, const_ps = pick(code_pc), pick(const_pc), pick(code_ps), pick(const_ps)
        self.nslots = int(len(first))
        cell_slot = np.full(nx * ny, -1, dtype=np.int32)
        if self.ncells:
            cell_slot[slots] = inverse.astype(np.int32)
        self.cell_slot = torch.as_tensor(cell_slot, device=device)
        self.code_pc = None if code_pc is None else torch.as_tensor(code_pc.astype(np.int8).ravel(), device=device)
        self.code_ps = None if code_ps is None else torch.as_tensor(code_ps.astype(np.int8).ravel(), device=device)
        self.const_pc = None if const_pc is None else torch.as_tensor(const_pc.ravel(), dtype=dtype, device=device)
        self.const_ps = None if const_ps is None else torch.as_tensor(const_ps.ravel(), dtype=dtype, device=device)
        # host copies for inspection and tests (rows per slot; cell_slot maps cells to slots)
        self.host = dict(cells=np.asarray(slots), cell_slot=cell_slot, code_pc=code_pc, const_pc=const_pc,
                         code_ps=code_ps, const_ps=const_ps)
        self.struct = nat.BCTable(self.nslots, nat.ptr(self.cell_slot) if self.nslots else None,
                                  nat.ptr(self.code_pc), nat.ptr(self.code_ps),
                                  nat.ptr(self.const_pc), nat.ptr(self.const_ps))

        self._ptr = ctypes.pointer(self.struct)

    def ref(self):
        """``xb_bc_table*`` for the C ABI (None when no cell carries a rule)."""
        return self._ptr if self.ncells else None


def compile_rules(bcs, nx, ny, q, stages=("PostCollision", "PostStreaming"), bounds=None):
    """Fold a BC list into per-cell rules.

    Returns ``(cells, {stage: (code, const)})`` with ``cells`` the sorted flat cell
    ids ``x*ny + y`` of every cell any BC writes.  ``bounds=(nxb, nyb)`` keeps only
    cells inside a smaller array (JAX drops out-of-range scatter rows)."""
    per_bc = []
    for bc in bcs:
        if bc.implementationStep not in stages:
            continue
        ix = np.asarray(bc.indices[0], dtype=np.int64)
        iy = np.asarray(bc.indices[1], dtype=np.int64)
        per_bc.append((bc, ix, iy))
    cells = set()
    for bc, ix, iy in per_bc:
        keep = (ix < (bounds[0] if bounds else nx)) & (iy < (bounds[1] if bounds else ny))
        cells.update((ix[keep] * ny + iy[keep]).tolist())
    cells = np.array(sorted(cells), dtype=np.int64)
    slot_of = {c: s for s, c in enumerate(cells.tolist())}
    out = {}
    for stage in stages:
        code = np.zeros((len(cells), q), dtype=np.int16)
        const = np.zeros((len(cells), q), dtype=np.float64)
        for bc, ix, iy in per_bc:
            if bc.implementationStep != stage:
                continue
            flat = ix * ny + iy
            keep = (ix < (bounds[0] if bounds else nx)) & (iy < (bounds[1] if bounds else ny))
            rows = np.array([slot_of[c] for c in flat[keep].tolist()], dtype=np.int64)
            # rows outside ``bounds`` were gathered at clamped indices by JAX and their
            # scatter dropped: they take part in the symbolic apply but are not written back
            full_code = np.zeros((len(ix), q), dtype=np.int16)
            full_const = np.zeros((len(ix), q), dtype=np.float64)
            full_code[keep] = code[rows]
            full_const[keep] = const[rows]
            st = RuleState(full_code, full_const)
            bc.apply_rules(st)
            code[rows] = st.code[keep]
            const[rows] = st.const[keep]
        out[stage] = (code, const)
    return cells, out


def build_rule_table(bcs, nx, ny, q, dtype, device, stages=("PostCollision", "PostStreaming"), bounds=None):
    """Rule table for the kernels, or None when no cell carries a rule."""
    cells, rules = compile_rules(bcs, nx, ny, q, stages, bounds)
    if len(cells) == 0:
        return None
    has_pc = "PostCollision" in rules and any(b.implementationStep == "PostCollision" for b in bcs)
    has_ps = "PostStreaming" in rules and any(b.implementationStep == "PostStreaming" for b in bcs)
    code_pc, const_pc = rules["PostCollision"] if has_pc else (None, None)
    code_ps, const_ps = rules["PostStreaming"] if has_ps else (None, None)
    if code_ps is None and "PostStreaming" in stages:
        code_ps = np.zeros((len(cells), q), dtype=np.int16)
        const_ps = np.zeros((len(cells), q))
    return RuleTable(nx, ny, q, cells, code_pc, const_pc, code_ps, const_ps, dtype, device)
