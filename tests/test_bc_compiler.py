"""CPU checks of the host BC compiler: the per-cell rules the kernels evaluate must
reproduce the reference's gather/scatter boundary passes (restated in the oracle)
on random populations, for every BC type, wall combination and multigrid level."""
import numpy as np
import pytest
import torch

import problems as P
from oracle import bc as obc
from src.boundary_conditions import compile_rules

ZERO = 100


def eval_rules(code, const, default, src):
    """Evaluate rule rows: value = base + const, base in {default, +/-src[k], 0}."""
    out = np.empty_like(default)
    for s in range(code.shape[0]):
        for d in range(code.shape[1]):
            c = int(code[s, d])
            if c == 0:
                base = default[s, d]
            elif c == ZERO:
                base = 0.0
            else:
                base = np.sign(c) * src[s, abs(c) - 1]
            out[s, d] = base + const[s, d]
    return out


def apply_compiled(bcs, fout, fin, stage):
    nx, ny, q = fout.shape
    cells, rules = compile_rules(bcs, nx, ny, q)
    out = fout.copy()
    if len(cells) == 0:
        return out
    code, const = rules[stage]
    x, y = cells // ny, cells % ny
    out[x, y] = eval_rules(code, const, fout[x, y], fin[x, y])
    return out


def _pairs():
    yield "neumann", lambda k: P.neumann_elasticity(k, n=8, levels=3)
    yield "dirichlet", lambda k: P.dirichlet_elasticity(k, n=8, levels=3)
    yield "mixed", lambda k: P.mixed_halfway_elasticity(k)
    yield "consolidation", lambda k: P.biot_consolidation(k, nx=16, levels=3)
    yield "darcy_c2", lambda k: P.darcy_c2(k, n=10)
    yield "darcy_eq", lambda k: P.darcy_equilibrium(k, n=8)
    yield "darcy_obstacle", lambda k: P.darcy_obstacle(k)
    yield "elastic_obstacle", lambda k: P.obstacle_elasticity(k)
    yield "elastic_obstacle_mg", lambda k: P.obstacle_elasticity(k, levels=3)


@pytest.mark.parametrize("name,make", list(_pairs()))
def test_rules_match_reference_scatter(name, make):
    g, o = make("gpu"), make("oracle")
    solvers = [(g.elasticity, o.elasticity), (g.darcy, o.darcy)] if hasattr(g, "elasticity") else [(g, o)]
    rng = np.random.default_rng(5)
    checked = 0
    for gs, os_ in solvers:
        levels = getattr(gs, "levels", 1)
        for level in range(levels):
            gb = gs.BCs if level == 0 else gs.BCs_MG[level]
            ob = os_.BCs if level == 0 else os_.BCs_MG[level]
            nx, ny = gs.grid_size(level) if hasattr(gs, "grid_size") else (gs.nx, gs.ny)
            q = gs.q
            for stage in ("PostCollision", "PostStreaming"):
                fout = rng.standard_normal((nx, ny, q))
                fin = rng.standard_normal((nx, ny, q))
                ref = obc.apply_bc(ob, torch.as_tensor(fout), torch.as_tensor(fin), stage).numpy()
                got = apply_compiled(gb, fout, fin, stage)
                np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)
                checked += 1
    assert checked > 0


def test_duplicate_index_quirk_is_kept():
    # halfway Dirichlet next to the y-halo of a periodic face: column 0 of the
    # reference's imissing is written by every non-missing direction, the last
    # write (post-collision f_0 of the same cell) wins.
    g = P.mixed_halfway_elasticity("gpu")
    left = g.BCs[0]
    cells, rules = compile_rules([left], g.nx, g.ny, 8)
    code, _ = rules["PostStreaming"]
    assert (code[:, 0] == 1).all()


def test_darcy_antibounceback_negates_rest_population():
    # ADR anti-bounce-back: fbd.at[bindex, imissing].set(-fin[bindex, iknown]) also
    # writes -fin[0] into the rest population (imissing = iknown = 0 off the wall)
    g = P.darcy_c2("gpu", n=6)
    cells, rules = compile_rules([g.BCs[0]], g.nx, g.ny, 9)
    code, const = rules["PostStreaming"]
    assert (code[:, 0] == -1).all()
    w0 = 4.0 / 9.0
    nonmissing = (np.asarray(g.BCs[0].imissing) == 0).sum(axis=1)
    np.testing.assert_allclose(const[:, 0], nonmissing * 2 * w0 * 1.0)


def test_coarse_gradient_rules_drop_out_of_range_rows():
    g = P.biot_consolidation("gpu", nx=16, levels=3)
    da = g.darcy
    bcs = [bc for bc in da.BCs if bc.name in ("ADR_Neumann_BounceBack", "ADR_AntiBounceBack")]
    nxc, nyc = 8, 10
    cells, _ = compile_rules(bcs, da.nx, da.ny, 9, stages=("PostStreaming",), bounds=(nxc, nyc))
    x, y = cells // da.ny, cells % da.ny
    assert (x < nxc).all() and (y < nyc).all()
    # the bottom wall (y = 1) survives, the top wall (y = ny - 2 = 16) is dropped
    assert set(y.tolist()) == {1}
