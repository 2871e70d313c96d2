"""Adjoint of the multigrid V-cycle and of the coupled Biot step, FP64.

Pins: a transposed V-cycle is the exact transpose of the V-cycle's linear part
(in the populations and in the static force), and its Lame-parameter derivatives
match central differences; the gradient of a displacement / pressure misfit
through checkpointed coupled steps matches ``torch.autograd`` through the oracle
restatement (the stand-in for ``jax.grad`` through the reference, which cannot
run here) and central finite differences in 16 random directions."""
import numpy as np
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    lam = 2.0 + 0.3 * rng.standard_normal((nx, ny))
    mu = 1.0 + 0.2 * rng.standard_normal((nx, ny))
    K = 0.5 * np.exp(0.5 * rng.standard_normal((nx, ny)))
    return lam, mu, 1.0 / (3.0 * K + 0.5)


def _problem(nx, levels, k, n_inner, seed=1, alt=False):
    p = P.biot_consolidation("gpu", nx=nx, levels=levels, steps=n_inner, k_smooth=k)
    p.alternative_gradient = alt
    lam, mu, om = _fields(nx, nx + 2, seed)
    p.elasticity.set_lame_fields(lam, mu)
    p.darcy.set_omega_field(om)
    return p


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("nx,levels,k", [(16, 2, 2), (16, 3, 2), (1024, 10, 8)])
def test_vcycle_adjoint_is_the_transpose(nx, levels, k):
    """At 16 x 18 with the Lame-parameter derivatives against central differences; at
    the C4 size (1024 x 1026, 10 levels, k = 8) the exact transposes only."""
    from src.biot_adjoint import BiotAdjointSession
    say("transposed V-cycle, consolidation", nx, "x", nx + 2, "levels", levels, "k", k)
    p = _problem(nx, levels, k, 3)
    s = BiotAdjointSession(p, 1, 4)
    el = p.elasticity
    nx, ny = el.nx, el.ny
    g = torch.Generator().manual_seed(0)

    def rnd(*shape, scale=1.0):
        return (scale * torch.randn(shape, generator=g, dtype=torch.float64)).cuda()
    f, d = rnd(8, nx, ny, scale=1e-3), rnd(8, nx, ny)
    base, dF = rnd(2, nx, ny, scale=1e-3), rnd(2, nx, ny)
    a = rnd(8, nx, ny)

    def V(f, base):
        out, _, states = s._vcycle_record(f, el._divergence_soa(f), base)
        return out, states
    out0, states = V(f, base)
    acc = {"lam": torch.zeros((1, nx, ny), dtype=torch.float64, device="cuda"),
           "mu": torch.zeros((1, nx, ny), dtype=torch.float64, device="cuda"),
           "base": torch.zeros((2, nx, ny), dtype=torch.float64, device="cuda")}
    ab = s._level_adjoint(0, a, states, None, acc)
    lhs = ((V(f + d, base)[0] - out0) * a).sum().item()
    rhs = (d * ab).sum().item()
    say("populations", lhs, rhs)
    assert _rel(lhs, rhs) < 1e-10
    lhs = ((V(f, base + dF)[0] - out0) * a).sum().item()
    rhs = (dF * acc["base"]).sum().item()
    say("static force", lhs, rhs)
    assert _rel(lhs, rhs) < 1e-10
    if nx > 64:
        return
    lam0, mu0 = el._lam_field.clone().reshape(nx, ny), el._mu_field.clone().reshape(nx, ny)
    h = 1e-6
    for name in ("lam", "mu"):
        dp = rnd(nx, ny)
        outs = []
        for sgn in (1.0, -1.0):
            if name == "lam":
                el.set_lame_fields(lam0 + sgn * h * dp, mu0)
            else:
                el.set_lame_fields(lam0, mu0 + sgn * h * dp)
            outs.append(V(f, base)[0])
        el.set_lame_fields(lam0, mu0)
        fd = ((outs[0] - outs[1]) * a).sum().item() / (2 * h)
        an = (dp * acc[name].reshape(nx, ny)).sum().item()
        say(name, fd, an)
        assert _rel(fd, an) < 1e-6


def _misfit_targets(nx, ny, seed):
    g = torch.Generator().manual_seed(seed)
    return (1e-4 * torch.randn((2, nx, ny), generator=g, dtype=torch.float64),
            1e-4 * torch.randn((1, nx, ny), generator=g, dtype=torch.float64))


def _misfit(eta, p, targets):
    te, tp = targets
    return 0.5 * ((eta - te.to(eta.device)) ** 2).sum() + 0.5 * ((p - tp.to(p.device)) ** 2).sum()


@pytest.mark.parametrize("alt", [False, True])
def test_biot_gradient_matches_autograd_through_oracle(alt):
    from src.biot_adjoint import biot_solve
    # three physical steps: a Darcy step from the zero state conserves sum f for any
    # omega, so the permeability enters the pressure from the second step on
    nx, levels, k, n_inner, steps = 16, 2, 2, 3, 3
    ny = nx + 2
    say("coupled gradient vs autograd through the oracle, alternative_gradient", alt)
    lam, mu, om = _fields(nx, ny, 2)
    targets = _misfit_targets(nx, ny, 3)
    p = P.biot_consolidation("gpu", nx=nx, levels=levels, steps=n_inner, k_smooth=k)
    p.alternative_gradient = alt
    # the recompute takes each V-cycle's coarse correction from the forward (alt False)
    # or solves the coarse levels again (alt True: no budget for keeping them)
    p.adjoint_coarse_budget_bytes = 0 if alt else 8e9
    ts = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for v in (om, lam, mu)]
    # K = 2 < 3 V-cycles per step: V-cycle checkpoints inside each step; K = 6:
    # checkpoints every second step, segments of two steps recomputed
    eta, pr = biot_solve(p, *ts, steps, checkpoint_every=6 if alt else 2)
    J = _misfit(eta, pr, targets)
    J.backward()
    to = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (om, lam, mu)]
    po = P.biot_consolidation("oracle", nx=nx, levels=levels, steps=n_inner, k_smooth=k, lambda_field=to[1],
                              mu_field=to[2], omega_field=to[0][..., None])
    po.alternative_gradient = alt
    fe, fd, dde = po.initial_state()
    for t in range(steps):
        fe, fd, eta_o, _, dde, p_o, _ = po.step(fe, fd, dde, t)
    Jo = _misfit(eta_o.permute(2, 0, 1), p_o.permute(2, 0, 1), targets)
    Jo.backward()
    say("J", J.item(), Jo.item())
    assert _rel(J.item(), Jo.item()) < 1e-12
    for name, a, b in zip(("omega", "lambda", "mu"), ts, to):
        ga, gb = a.grad.cpu(), b.grad
        err = (ga - gb).abs().max().item() / gb.abs().max().item()
        say(name, "max rel err", err, "norm", gb.norm().item())
        assert err < 1e-9, name


def test_biot_gradient_matches_finite_differences_16_directions():
    from src.biot_adjoint import biot_solve
    nx, levels, k, n_inner, steps = 16, 2, 2, 4, 3
    ny = nx + 2
    say("coupled gradient vs central differences, 16 directions")
    lam, mu, om = _fields(nx, ny, 4)
    targets = _misfit_targets(nx, ny, 5)
    p = P.biot_consolidation("gpu", nx=nx, levels=levels, steps=n_inner, k_smooth=k)
    ts = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for v in (om, lam, mu)]
    eta, pr = biot_solve(p, *ts, steps, checkpoint_every=3)
    _misfit(eta, pr, targets).backward()
    grads = [t.grad.clone() for t in ts]
    base = [t.detach().clone() for t in ts]
    g = torch.Generator().manual_seed(6)
    h = 1e-5
    worst = 0.0
    with torch.no_grad():
        for i in range(16):
            ds = [(b.abs() * torch.randn(b.shape, generator=g, dtype=torch.float64).cuda()) for b in base]
            Js = []
            for sgn in (1.0, -1.0):
                args = [b + sgn * h * d for b, d in zip(base, ds)]
                e, q = biot_solve(p, *args, steps, checkpoint_every=3)
                Js.append(_misfit(e, q, targets).item())
            fd = (Js[0] - Js[1]) / (2 * h)
            an = sum((gr * d).sum().item() for gr, d in zip(grads, ds))
            worst = max(worst, _rel(fd, an))
            say(i, fd, an)
    say("worst relative difference", worst)
    assert worst < 1e-6


def test_c4_geometry_gradient_matches_finite_differences_at_128():
    # the C4 set-up at 128 x 130, FP64: 7 levels by the rule of
    # examples/poroelasticity/2D.py:216-219, 8 smoothing steps, 2 V-cycles per step,
    # log-normal K (variance 2, correlation 32, seed 1), lambda = 2 and mu = 1 held as
    # fields, target from K perturbed by a log-normal factor (seed 2); gradient of the
    # misfit with respect to K (through omega = 1 / (3 K + 1/2)) and mu along 16
    # random directions against central differences; checkpoints every 4 V-cycles
    from src.biot_adjoint import biot_solve
    n, steps = 128, 6
    levels = n.bit_length() - 1
    while n % 2 ** (levels - 1):
        levels -= 1
    say("C4 geometry gradient at", n, "levels", levels)
    p = P.biot_consolidation("gpu", nx=n, levels=levels, steps=2, k_smooth=8)
    nx, ny = p.elasticity.nx, p.elasticity.ny
    Kf = 0.5 * torch.as_tensor(P.lognormal_field(nx, ny, 2.0, 32.0, 1), dtype=torch.float64).cuda()
    Kt = Kf * torch.as_tensor(P.lognormal_field(nx, ny, 0.1, 32.0, 2), dtype=torch.float64).cuda()
    lam = torch.full((nx, ny), 2.0, dtype=torch.float64, device="cuda")
    mu = torch.full((nx, ny), 1.0, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        te, tp = (t.clone() for t in biot_solve(p, 1.0 / (3.0 * Kt + 0.5), lam, mu, steps, checkpoint_every=4))

    def J(K, m):
        e, q = biot_solve(p, 1.0 / (3.0 * K + 0.5), lam, m, steps, checkpoint_every=4)
        return 0.5 * ((e - te) ** 2).sum() + 0.5 * ((q - tp) ** 2).sum()

    Kg, mg = Kf.clone().requires_grad_(True), mu.clone().requires_grad_(True)
    J(Kg, mg).backward()
    gK, gm = Kg.grad.clone(), mg.grad.clone()
    assert torch.isfinite(gK).all() and torch.isfinite(gm).all() and gK.abs().max() > 0
    g = torch.Generator().manual_seed(7)
    h = 1e-5
    worst = 0.0
    with torch.no_grad():
        for i in range(16):
            dK = Kf * torch.randn((nx, ny), generator=g, dtype=torch.float64).cuda()
            dm = mu * torch.randn((nx, ny), generator=g, dtype=torch.float64).cuda()
            fd = (J(Kf + h * dK, mu + h * dm).item() - J(Kf - h * dK, mu - h * dm).item()) / (2 * h)
            an = (gK * dK).sum().item() + (gm * dm).sum().item()
            worst = max(worst, _rel(fd, an))
            say(i, fd, an)
    say("worst relative difference", worst)
    assert worst < 1e-6


@pytest.mark.parametrize("case,precision", [("consolidation", "f64/f64"), ("consolidation", "f32/f32"),
                                            ("periodic", "f64/f64")])
@pytest.mark.parametrize("level", [1, 2])
def test_small_grid_transposed_smoothing_matches_sweeps(case, precision, level):
    """xb_elast_smooth_adjoint_small (a run of transposed smoothing sweeps with their
    transposed coupling force in one workgroup) and xb_elast_smooth_adjoint_tiles
    (1, 2 or 3 of them per launch on LDS tiles) are bitwise the sequence of
    xb_elast_mg_adjoint and xb_elast_coupling_adjoint calls they replace."""
    from src import _native as nat
    from src.biot_adjoint import BiotAdjointSession
    say("small-grid transposed smoothing", case, precision, "level", level)
    if case == "consolidation":
        p = P.biot_consolidation("gpu", nx=32, levels=3, steps=2, k_smooth=4, precision=precision)
    else:
        p = P.biot_periodic("gpu", nx=32, levels=3, steps=2, precision=precision)
    s = BiotAdjointSession(p, 1, 4)
    el = p.elasticity
    nx, ny = el.grid_size(level)
    x0, x1, y0, y1 = el.interior_box(level)
    g = torch.Generator().manual_seed(7)
    ab = torch.randn((8, nx, ny), generator=g, dtype=torch.float64).to(el._tdt).cuda()
    rbar0 = torch.randn((8, x1 - x0, y1 - y0), generator=g, dtype=torch.float64).to(el._tdt).cuda()
    no_acc = {"lam": None, "mu": None, "base": None}
    for sweeps in (1, 3, el.k_smooth):
        rb1, a = rbar0.clone(), ab
        for _ in range(sweeps):
            a = s._sweep_adjoint(level, nat.XB_ADJ_SMOOTH, a, None, None, rb1, no_acc)
        rb2 = rbar0.clone()
        b = s._smooth_adjoint_small(level, ab, sweeps, rb2)
        assert torch.equal(a, b), (sweeps, (a - b).abs().max().item())
        assert torch.equal(rb1, rb2), sweeps
        if sweeps == el.k_smooth:
            # the same run on LDS tiles (xb_elast_smooth_adjoint_tiles), 1 to 3 sweeps
            # per launch
            for m in (1, 2, 3):
                s.adjoint_tile_steps = m
                rb3 = rbar0.clone()
                c = s._smooth_adjoint_run(level, ab, [None] * sweeps, rb3, no_acc)
                assert torch.equal(a, c), (m, (a - c).abs().max().item())
                assert torch.equal(rb1, rb3), m


@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
def test_tiled_transposed_smoothing_matches_sweep_pairs(precision):
    """The gradient with the transposed smoothing on LDS tiles (2 or 3 sweeps per
    launch, xb_elast_smooth_adjoint_tiles) equals, bit for bit, the gradient with one
    xb_elast_mg_adjoint + xb_elast_coupling_adjoint pair per sweep: consolidation
    walls, Darcy wall rules of the coupling gradient, Lame fields, parameter
    derivatives on the finest level, residual adjoints on a coarse level of 32 x 34."""
    from src.biot_adjoint import biot_solve
    nx, levels, k, n_inner, steps = 64, 3, 4, 2, 2
    ny = nx + 2
    dt = torch.float64 if precision == "f64/f64" else torch.float32
    lam, mu, om = _fields(nx, ny, 7)
    te, tp = _misfit_targets(nx, ny, 8)
    grads, Js = {}, {}
    for m in (0, 2, 3, None):
        # None: the default, 2 sweeps per launch on levels 0 and 1 and 3 on level 2
        say("transposed smoothing sweeps per launch", m, precision)
        p = P.biot_consolidation("gpu", nx=nx, levels=levels, steps=n_inner, k_smooth=k, precision=precision)
        if m is not None:
            p.adjoint_tile_steps = m
        ts = [torch.tensor(v, dtype=dt, device="cuda", requires_grad=True) for v in (om, lam, mu)]
        eta, pr = biot_solve(p, *ts, steps, checkpoint_every=2)
        J = 0.5 * ((eta - te.to(eta)) ** 2).sum(dtype=torch.float64) + \
            0.5 * ((pr - tp.to(pr)) ** 2).sum(dtype=torch.float64)
        J.backward()
        Js[m] = J.item()
        grads[m] = [t.grad.clone() for t in ts]
        say("J", Js[m], "grad norms", [g.double().norm().item() for g in grads[m]])
    for m in (2, 3, None):
        assert Js[m] == Js[0]
        for name, a, b in zip(("omega", "lambda", "mu"), grads[0], grads[m]):
            assert torch.equal(a, b), (name, m, (a - b).abs().max().item())
