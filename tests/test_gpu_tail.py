"""The single-workgroup coarse tail of the V-cycle (xb_vcycle_tail), in its LDS and its
global-memory form, against the level-by-level launches: bit-identical states and
residual norms, with and without walls and the coupling force, FP64 and FP32."""
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _cycles(vg, f0, n):
    vg.load(f0)
    norms = []
    for _ in range(n):
        vg.replay()
        norms.append(vg.norm.item())
    return vg.state.clone(), norms


def _compare(make, lds, n=4, seed=0):
    outs = []
    for use_tail in (False, True):
        el, ext, nlf = make()
        el.use_tail = use_tail
        el.tail_in_lds = lds
        g = torch.Generator().manual_seed(seed)
        f0 = (1e-3 * torch.randn((8, el.nx, el.ny), generator=g, dtype=torch.float64)).to(el._tdt).cuda()
        vg = el.vcycle_graph(ext, nlf)
        outs.append(_cycles(vg, f0, n))
        if use_tail:
            start = el._tail_start()
            say("tail starts at level", start, "of", el.levels, "grids",
                [el.grid_size(l) for l in range(el.levels)])
            assert start is not None
            tails = [t for k, t in el.__dict__.get("_tails", {}).items() if k[0] == start]
            assert tails and all((t.plan.lds_bytes > 0) == lds for t in tails)
    (a, na), (b, nb) = outs
    assert torch.equal(a, b), (a - b).abs().max().item()
    assert na == nb


@pytest.mark.parametrize("lds", [True, False])
@pytest.mark.parametrize("kind,n,levels", [("neumann", 32, 4), ("dirichlet", 32, 3)])
def test_tail_matches_levelwise_with_walls(kind, n, levels, lds):
    say("tail vs level-wise launches,", kind, n, levels)

    def make():
        el = (P.neumann_elasticity if kind == "neumann" else P.dirichlet_elasticity)("gpu", n=n, levels=levels)
        return el, None, None
    _compare(make, lds)


@pytest.mark.parametrize("lds", [True, False])
@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
def test_tail_matches_levelwise_periodic(precision, lds):
    say("tail vs level-wise launches, periodic 64, 4 levels", precision)

    def make():
        return P.periodic_elasticity("gpu", nx=64, levels=4, precision=precision), None, None
    _compare(make, lds)


@pytest.mark.parametrize("lds", [True, False])
@pytest.mark.parametrize("levels", [3, 5])
def test_tail_matches_levelwise_coupled(levels, lds):
    say("tail vs level-wise launches, consolidation coupling force, levels", levels)

    def make():
        p = P.biot_consolidation("gpu", nx=32, levels=levels, k_smooth=4)
        el = p.elasticity
        g = torch.Generator().manual_seed(3)
        ext = (1e-4 * torch.randn((2, el.nx, el.ny), generator=g, dtype=torch.float64)).cuda()
        return el, ext, p.coupling_force()
    _compare(make, lds)
