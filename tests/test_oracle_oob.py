"""The oracle's out-of-range index handling in the Darcy gradient on coarse multigrid
levels (reference src/darcy.py:75-104 called through poroelasticity.py:84-86 on every
level of the elasticity V-cycle, src/elasticity.py:1221-1312): the Darcy solver's wall
indices belong to the finest grid, so on a coarse array some of them lie outside it.
JAX clamps such indices in a gather and drops such updates in a scatter ("JAX - The
Sharp Bits", out-of-bounds indexing); the oracle does the same (oracle/darcy.py
``_inbounds`` / ``_apply_clamped``).

These tests show that the clamp cannot matter: the ADR wall rules are row-local (a
rule row reads the populations of its own cell and its own constants only), so the
value gathered for a dropped row never reaches a kept one.  Pointing the out-of-range
rows at a NaN-filled cell or at wrapped indices instead of clamped ones gives the same
gradient bit for bit (and a finite one).  The one convention the result depends on is
the scatter's drop, the documented default of JAX's ``.at[].set``."""
import numpy as np
import pytest
import torch

import problems as P


def _coarse_inputs(nx, levels, seed):
    o = P.biot_consolidation("oracle", nx=nx, levels=levels, steps=1)
    el, da = o.elasticity, o.darcy
    g = torch.Generator().manual_seed(seed)
    out = []
    for level in range(1, levels):
        cx, cy = el.grid_size(level)
        out.append(torch.randn((cx, cy, 1), generator=g, dtype=torch.float64))
    return da, out


def _gradient_with(da, q, redirect):
    """darcy.gradient(q) with every out-of-range wall row of the gather sent to
    ``redirect(ix, iy, shape)`` (indices into an array padded by one NaN row and
    column) instead of clamped, then dropped as in the oracle."""
    from oracle.bc import stream
    p_in = q.repeat(1, 1, da.q)
    p_out = stream(p_in, da.lattice)
    nx, ny = q.shape[0], q.shape[1]
    for bc in da.BCs:
        if bc.name not in ("ADR_Neumann_BounceBack", "ADR_AntiBounceBack"):
            continue
        ix = np.asarray(bc.indices[0])
        iy = np.asarray(bc.indices[1])
        keep = (ix < nx) & (iy < ny)
        pad_out = torch.full((nx + 1, ny + 1, da.q), float("nan"), dtype=p_out.dtype)
        pad_in = pad_out.clone()
        pad_out[:nx, :ny] = p_out
        pad_in[:nx, :ny] = p_in
        rx, ry = redirect(ix, iy, (nx, ny))
        saved = bc.indices
        bc.indices = (np.where(keep, ix, rx), np.where(keep, iy, ry))
        try:
            vals = bc.apply(pad_out, pad_in)
        finally:
            bc.indices = saved
        p_out = p_out.clone()
        kx = torch.as_tensor(ix[keep], dtype=torch.long)
        ky = torch.as_tensor(iy[keep], dtype=torch.long)
        p_out[kx, ky] = vals[torch.as_tensor(keep)].to(p_out.dtype)
    w = torch.as_tensor(da.w, dtype=q.dtype)
    c = torch.as_tensor(da.c, dtype=q.dtype)
    return torch.matmul(p_out, -3 * w[..., None] * c.T)


@pytest.mark.parametrize("nx,levels", [(16, 3), (24, 4), (32, 3)])
def test_coarse_gradient_does_not_depend_on_the_gather_mode(nx, levels):
    da, qs = _coarse_inputs(nx, levels, seed=nx + levels)
    modes = {
        "nan": lambda ix, iy, s: (np.full_like(ix, s[0]), np.full_like(iy, s[1])),
        "wrap": lambda ix, iy, s: (ix % s[0], iy % s[1]),
        "zero": lambda ix, iy, s: (np.zeros_like(ix), np.zeros_like(iy)),
    }
    some_out_of_range = False
    for q in qs:
        ref = da.gradient(q)  # clamp on gather, drop on scatter
        assert torch.isfinite(ref).all()
        for bc in da.BCs:
            ix, iy = np.asarray(bc.indices[0]), np.asarray(bc.indices[1])
            some_out_of_range |= bool(((ix >= q.shape[0]) | (iy >= q.shape[1])).any())
        for name, redirect in modes.items():
            got = _gradient_with(da, q, redirect)
            assert torch.isfinite(got).all(), name
            assert torch.equal(got, ref), (name, (got - ref).abs().max().item())
    assert some_out_of_range, "the case must hold wall rows outside the coarse grid"


def test_dropped_rows_leave_their_cells_as_streamed():
    """Cells of a coarse grid that no in-range wall row names keep the streamed values:
    the gradient equals the plain streamed stencil there (the scatter writes kept rows
    only)."""
    from oracle.bc import stream
    da, qs = _coarse_inputs(16, 3, seed=7)
    q = qs[0]
    nx, ny = q.shape[0], q.shape[1]
    named = np.zeros((nx, ny), dtype=bool)
    for bc in da.BCs:
        if bc.name in ("ADR_Neumann_BounceBack", "ADR_AntiBounceBack"):
            ix, iy = np.asarray(bc.indices[0]), np.asarray(bc.indices[1])
            keep = (ix < nx) & (iy < ny)
            named[ix[keep], iy[keep]] = True
    p_out = stream(q.repeat(1, 1, da.q), da.lattice)
    w = torch.as_tensor(da.w, dtype=q.dtype)
    c = torch.as_tensor(da.c, dtype=q.dtype)
    plain = torch.matmul(p_out, -3 * w[..., None] * c.T)
    ref = da.gradient(q)
    free = torch.as_tensor(~named)
    assert free.any() and (~free).any()
    assert torch.equal(ref[free], plain[free])
