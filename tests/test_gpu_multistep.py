"""Several plain steps per launch on LDS tiles (xb_darcy_stepn / xb_elast_stepn)
against the same number of single-step launches: bit-identical populations, with
walls, per-cell parameters, forces and sources, on grids that are and are not
multiples of the tile, FP64 and FP32."""
import numpy as np
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _single(step, f, n):
    a = f.clone()
    b = torch.empty_like(a)
    for _ in range(n):
        step(a, b)
        a, b = b, a
    return a


@pytest.mark.parametrize("n,precision,hetero", [(514, "f64/f64", True), (70, "f32/f32", True), (16, "f64/f64", False),
                                               (5, "f64/f64", False)])
@pytest.mark.parametrize("m", [1, 3, 4, 6, 8])
@pytest.mark.parametrize("threads", [256, 512, 704])
def test_darcy_stepn_matches_single_steps(n, precision, hetero, m, threads):
    say("darcy", n, precision, "hetero", hetero, "steps per launch", m, "threads per tile", threads)
    d = P.darcy_c2("gpu", n=n - 2, precision=precision, heterogeneous=hetero, seed=0, variance=1.0, corr=4.0)
    g = torch.Generator().manual_seed(1)
    f = (1.0 + 0.1 * torch.randn((9, d.nx, d.ny), generator=g, dtype=torch.float64)).to(d._tdt).cuda()
    src = (1e-3 * torch.randn((1, d.nx, d.ny), generator=g, dtype=torch.float64)).to(d._tdt).cuda()
    ref = _single(lambda a, b: d._step_soa(a, src, False, out=b), f, m)
    got = d._stepn_into(f, torch.empty_like(f), m, src, threads=threads)
    assert torch.equal(got, ref), (got - ref).abs().max().item()


@pytest.mark.parametrize("case", ["periodic", "walls", "fields"])
@pytest.mark.parametrize("m", [2, 5, 8])
@pytest.mark.parametrize("threads", [256, 512, 704])
def test_elast_stepn_matches_single_steps(case, m, threads):
    say("elasticity", case, "steps per launch", m, "threads per tile", threads)
    g = torch.Generator().manual_seed(2)
    if case == "periodic":
        el = P.periodic_elasticity("gpu", nx=64)
    elif case == "walls":
        el = P.consolidation_elasticity("gpu", nx=40)
    else:
        rng = np.random.default_rng(3)
        el = P.consolidation_elasticity("gpu", nx=33, lambda_field=rng.uniform(1.0, 3.0, (33, 35)),
                                        mu_field=rng.uniform(0.5, 1.5, (33, 35)))
    f = (1e-3 * torch.randn((8, el.nx, el.ny), generator=g, dtype=torch.float64)).to(el._tdt).cuda()
    ext = (1e-4 * torch.randn((2, el.nx, el.ny), generator=g, dtype=torch.float64)).to(el._tdt).cuda()
    ref = _single(lambda a, b: el._step_into(a, b, ext), f, m)
    got = el._stepn_into(f, torch.empty_like(f), m, ext, threads)
    assert torch.equal(got, ref), (got - ref).abs().max().item()


def test_run_steps_with_tuned_multistep_matches_single_steps():
    say("darcy run_steps after tuning the steps per launch")
    d = P.darcy_c2("gpu", n=126, precision="f64/f64", heterogeneous=True, seed=0, variance=1.0, corr=4.0)
    d.autotune()
    choice, m, _ = d._multi()
    say("chosen launch kind", choice, "steps per launch", m)
    f0 = d._soa(d.assign_fields_sharded(), 9)
    ref = _single(lambda a, b: d._step_soa(a, None, False, out=b), f0, 77)
    got = d._soa(d.run_steps(f0.permute(1, 2, 0), 77, chunk=32), 9)
    assert torch.equal(got, ref)


def _smooth_cases(case, precision):
    import numpy as np
    g = torch.Generator().manual_seed(4)
    if case == "periodic_biot":
        p = P.biot_periodic("gpu", nx=96, levels=3, precision=precision)
        level = 0
    else:
        rng = np.random.default_rng(5)
        p = P.biot_consolidation("gpu", nx=64, levels=3, k_smooth=4, precision=precision)
        if case == "fields":
            p.elasticity.set_lame_fields(rng.uniform(1.5, 2.5, (64, 66)), rng.uniform(0.7, 1.3, (64, 66)))
        level = 1 if case == "coarse" else 0
    el = p.elasticity
    nx, ny = el.grid_size(level)
    x0, x1, y0, y1 = el.interior_box(level)
    dt = el._tdt
    f = (1e-3 * torch.randn((8, nx, ny), generator=g, dtype=torch.float64)).to(dt).cuda()
    r = (1e-4 * torch.randn((8, x1 - x0, y1 - y0), generator=g, dtype=torch.float64)).to(dt).cuda() \
        if level > 0 else None
    base = (1e-4 * torch.randn((2, nx, ny), generator=g, dtype=torch.float64)).to(dt).cuda() if level == 0 else None
    return el, p.coupling_force(), level, f, r, base


@pytest.mark.parametrize("case", ["periodic_biot", "consolidation", "fields", "coarse"])
@pytest.mark.parametrize("m,precision", [(2, "f64/f64"), (3, "f64/f64"), (4, "f64/f64"), (6, "f64/f64"),
                                         (2, "f32/f32"), (4, "f32/f32")])
def test_coupled_smoothn_matches_sweeps(case, m, precision):
    say("coupled smoothing on tiles", case, m, precision)
    el, nlf, level, f, r, base = _smooth_cases(case, precision)
    div = el._divergence_soa(f)
    a, da = f.clone(), div.clone()
    for _ in range(m):
        a, da = el._smooth_nl(a, da, level, r, base, nlf)
    b, db = el._smoothn_nl(f, div, level, r, base, nlf, m)
    assert torch.equal(a, b), (a - b).abs().max().item()
    assert torch.equal(da, db), (da - db).abs().max().item()


@pytest.mark.parametrize("case", ["periodic_biot", "consolidation", "fields", "coarse"])
@pytest.mark.parametrize("m,precision", [(2, "f64/f64"), (3, "f64/f64"), (4, "f64/f64"), (2, "f32/f32")])
def test_coupled_smoothn_with_coarse_correction_matches_prolong_then_sweeps(case, m, precision):
    """The post-smoothing launch with the prolongation fused in (coarse_out read by
    xb_elast_smoothn) against xb_elast_prolong_add_div followed by single sweeps;
    f_in is left as it was."""
    say("coupled smoothing on tiles from f + P c", case, m, precision)
    el, nlf, level, f, r, base = _smooth_cases(case, precision)
    g = torch.Generator().manual_seed(6)
    nxc, nyc = el.grid_size(level + 1)
    c = (1e-4 * torch.randn((8, nxc, nyc), generator=g, dtype=torch.float64)).to(f.dtype).cuda()
    div = el._divergence_soa(f)
    a, da = f.clone(), div.clone()
    el._prolong_add_div(a, da, c, level)
    for _ in range(m):
        a, da = el._smooth_nl(a, da, level, r, base, nlf)
    f0 = f.clone()
    b, db = el._smoothn_nl(f, div, level, r, base, nlf, m, coarse=c)
    assert torch.equal(f, f0)
    assert torch.equal(a, b), (a - b).abs().max().item()
    assert torch.equal(da, db), (da - db).abs().max().item()


@pytest.mark.parametrize("case", ["periodic_biot", "coarse"])
@pytest.mark.parametrize("m,precision", [(2, "f64/f64"), (4, "f64/f64"), (3, "f32/f32")])
def test_coupled_smoothn_from_zero_state_matches_zero_input(case, m, precision):
    """xb_elast_smoothn with a null f_in (a coarse level's zero start state, nothing
    read) equals the same launch from a zero-filled state, bit for bit."""
    say("coupled smoothing on tiles from the zero state", case, m, precision)
    el, nlf, level, f, r, base = _smooth_cases(case, precision)
    z = torch.zeros_like(f)
    a, da = el._smoothn_nl(z, el._divergence_soa(z), level, r, base, nlf, m)
    b, db = el._smoothn_nl(None, None, level, r, base, nlf, m)
    assert torch.equal(a, b), (a - b).abs().max().item()
    assert torch.equal(da, db), (da - db).abs().max().item()
