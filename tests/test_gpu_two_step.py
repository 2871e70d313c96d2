"""Two steps per launch (xb_elast_step2 / xb_darcy_step2) against two single steps.

Per cell both paths do the same arithmetic in the same order (the library is built
without floating-point contraction), so the populations must agree bit for bit:
with walls, per-cell parameters, sources and body forces, for strips that end
inside the grid and grids narrower than one strip."""
import numpy as np
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def _rand_like(t, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(t.shape, generator=g, dtype=torch.float64).to(t.dtype).cuda()


@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
@pytest.mark.parametrize("n", [14, 70, 133])
def test_darcy_two_steps_equal_two_single_steps(precision, n):
    d = P.darcy_c2("gpu", n=n, precision=precision, source=True)
    f = _rand_like(torch.empty((9, d.nx, d.ny), dtype=d._tdt), 1)
    a, b = torch.empty_like(f), torch.empty_like(f)
    d._step_soa(f, None, False, out=a)
    d._step_soa(a, None, False, out=b)
    c = d._step2_into(f, torch.empty_like(f))
    assert torch.equal(b, c), (b - c).abs().max().item()
    from src import _native as nat
    for rows in (2, 8, 64):
        out = torch.empty_like(f)
        args = d._args(f, out, None, None, True)
        args.rows = rows
        nat.check(nat.lib().xb_darcy_step2(args, nat.stream()), "xb_darcy_step2")
        assert torch.equal(b, out), rows


@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
@pytest.mark.parametrize("case", ["mixed", "consolidation_fields", "periodic"])
def test_elasticity_two_steps_equal_two_single_steps(precision, case):
    if case == "mixed":
        el = P.mixed_halfway_elasticity("gpu", nx=30, ny=67, precision=precision)
    elif case == "periodic":
        el = P.periodic_elasticity("gpu", nx=64, precision=precision)
    else:
        el = P.consolidation_elasticity("gpu", nx=40, precision=precision)
        rng = np.random.default_rng(3)
        el.set_lame_fields(rng.uniform(0.5, 2.0, (el.nx, el.ny)), rng.uniform(0.5, 2.0, (el.nx, el.ny)))
    f = 1e-3 * _rand_like(torch.empty((8, el.nx, el.ny), dtype=el._tdt), 2)
    ext = 1e-4 * _rand_like(torch.empty((2, el.nx, el.ny), dtype=el._tdt), 3)
    a, b = torch.empty_like(f), torch.empty_like(f)
    el._step_into(f, a, ext)
    el._step_into(a, b, ext)
    c = el._step2_into(f, torch.empty_like(f), ext)
    assert torch.equal(b, c), (b - c).abs().max().item()


def test_run_steps_uses_pairs_and_matches_single_steps():
    d = P.darcy_c2("gpu", n=30)
    f0 = d.assign_fields_sharded()
    got = d.run_steps(f0, 37, chunk=8)  # 4 graph replays of 4 two-step launches, 2 pairs, 1 single
    f = d._soa(f0, 9)
    for _ in range(37):
        f, _ = d._step_soa(f)
    assert torch.equal(got.permute(2, 0, 1), f)
