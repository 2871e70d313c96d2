"""Adjoint steps and gradients through the C ABI.

* The adjoint kernels are the transposes of the forward steps: for random state f,
  perturbation d and adjoint a, <S(f + d) - S(f), a> = <d, (dS/df)^T a> (the step is
  affine in f), with walls and per-cell parameters; their parameter derivatives agree
  with central differences of <S(f; theta), a>.
* Gradients of a misfit after n steps (checkpointed, graph-replayed backward) agree
  with torch.autograd through the same n steps of the CPU restatement of the
  reference (FP64, 1e-10), and with central finite differences along random
  directions (1e-6).
* The fixed-point adjoint of a converged steady state agrees with finite
  differences of re-converged steady states.
"""
import numpy as np
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _fields(nx, ny, seed, lo=0.6, hi=1.8):
    g = np.random.default_rng(seed)
    return g.uniform(lo, hi, (nx, ny)), g.uniform(lo, hi, (nx, ny))


def _cuda(a):
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device="cuda")


def _rand(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return (scale * torch.randn(shape, generator=g, dtype=torch.float64)).cuda()


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def _elastic(walls, nx=24):
    return P.consolidation_elasticity("gpu", nx=nx) if walls else P.periodic_elasticity("gpu", nx=nx)


@pytest.mark.parametrize("walls", [False, True])
def test_elasticity_adjoint_step_is_the_transpose(walls):
    from src.adjoint import ElasticityStepper
    say("elasticity adjoint transpose, walls =", walls)
    el = _elastic(walls)
    nx, ny = el.nx, el.ny
    lam, mu = (_cuda(t).reshape(1, nx, ny) for t in _fields(nx, ny, 1))
    force = _rand((2, nx, ny), 2, 1e-3)
    st = ElasticityStepper(el, lam, mu, force)
    f, d, a = _rand((8, nx, ny), 3), _rand((8, nx, ny), 4), _rand((8, nx, ny), 5)
    S = lambda x: st.step(x, torch.empty_like(x))
    lhs = _dot(S(f + d) - S(f), a)
    grads = st.grad_buffers(True)
    ata = st.adjoint(a, f, torch.empty_like(a), grads)
    rhs = _dot(d, ata)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)
    # parameter derivatives against central differences of <S(f; theta), a>
    eps = 1e-6
    for name in ("lam", "mu"):
        delta = _rand((1, nx, ny), 6 if name == "lam" else 7)
        vals = []
        for sgn in (1.0, -1.0):
            l2 = lam + sgn * eps * delta if name == "lam" else lam
            m2 = mu + sgn * eps * delta if name == "mu" else mu
            s2 = ElasticityStepper(el, l2, m2, force)
            vals.append(_dot(s2.step(f, torch.empty_like(f)), a))
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = _dot(grads[name], delta)
        assert abs(fd - ad) <= 1e-7 * abs(ad), (name, fd, ad)
    # the force enters affinely: one difference is exact
    dF = _rand((2, nx, ny), 8)
    s2 = ElasticityStepper(el, lam, mu, force + dF)
    fd = _dot(s2.step(f, torch.empty_like(f)) - S(f), a)
    assert abs(fd - _dot(grads["force"], dF)) <= 1e-11 * abs(fd)


@pytest.mark.parametrize("walls", [False, True])
def test_darcy_adjoint_step_is_the_transpose(walls):
    from src.adjoint import DarcyStepper
    say("darcy adjoint transpose, walls =", walls)
    da = P.darcy_c2("gpu", n=22) if walls else P.biot_periodic("gpu", nx=24).darcy
    nx, ny = da.nx, da.ny
    om = _cuda(np.random.default_rng(11).uniform(0.3, 1.7, (1, nx, ny)))
    src = _rand((1, nx, ny), 12, 1e-2)
    st = DarcyStepper(da, om, src)
    f, d, a = _rand((9, nx, ny), 13), _rand((9, nx, ny), 14), _rand((9, nx, ny), 15)
    S = lambda x: st.step(x, torch.empty_like(x))
    lhs = _dot(S(f + d) - S(f), a)
    grads = st.grad_buffers(True)
    rhs = _dot(d, st.adjoint(a, f, torch.empty_like(a), grads))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)
    eps = 1e-6
    delta = _rand((1, nx, ny), 16)
    vals = [_dot(DarcyStepper(da, om + s_ * eps * delta, src).step(f, torch.empty_like(f)), a) for s_ in (1.0, -1.0)]
    fd = (vals[0] - vals[1]) / (2 * eps)
    ad = _dot(grads["omega"], delta)
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    dS = _rand((1, nx, ny), 17)
    fd = _dot(DarcyStepper(da, om, src + dS).step(f, torch.empty_like(f)) - S(f), a)
    assert abs(fd - _dot(grads["source"], dS)) <= 1e-11 * abs(fd)


def _oracle_elastic_run(nx, lam_np, mu_np, n, target):
    lam_o = torch.tensor(lam_np, dtype=torch.float64, requires_grad=True)
    mu_o = torch.tensor(mu_np, dtype=torch.float64, requires_grad=True)
    o = P.consolidation_elasticity("oracle", nx=nx, lambda_field=lam_o, mu_field=mu_o)
    f = torch.zeros((o.nx, o.ny, 8), dtype=torch.float64)
    for k in range(n):
        f, _ = o._step(f, k)
    J = 0.5 * ((f - target) ** 2).sum()
    J.backward()
    return f, lam_o.grad, mu_o.grad


@pytest.mark.parametrize("use_graphs", [True, False])
def test_elasticity_gradient_matches_autograd_through_the_oracle(use_graphs):
    from src.adjoint import elasticity_solve
    nx, n, K = 16, 30, 8
    say("elasticity gradient vs oracle autograd", nx, n, K, use_graphs)
    el = P.consolidation_elasticity("gpu", nx=nx)
    ny = el.ny
    lam_np, mu_np = _fields(nx, ny, 21)
    target = 1e-3 * torch.randn((nx, ny, 8), generator=torch.Generator().manual_seed(22), dtype=torch.float64)
    fo, glo, gmo = _oracle_elastic_run(nx, lam_np, mu_np, n, target)
    lam = _cuda(lam_np).requires_grad_(True)
    mu = _cuda(mu_np).requires_grad_(True)
    f0 = torch.zeros((8, nx, ny), dtype=torch.float64, device="cuda")
    fN = elasticity_solve(el, lam, mu, f0, n, checkpoint_every=K, use_graphs=use_graphs)
    J = 0.5 * ((fN - target.permute(2, 0, 1).cuda()) ** 2).sum()
    J.backward()
    rel = lambda a, b: float((a.cpu() - b).abs().max() / b.abs().max())
    assert rel(fN.permute(1, 2, 0), fo.detach()) < 1e-12
    assert rel(lam.grad, glo) < 1e-10, rel(lam.grad, glo)
    assert rel(mu.grad, gmo) < 1e-10, rel(mu.grad, gmo)


def test_darcy_gradient_matches_autograd_through_the_oracle():
    from src.adjoint import darcy_solve
    n_steps, K = 24, 10
    say("darcy gradient vs oracle autograd")
    d = P.darcy_c2("gpu", n=14)
    nx, ny = d.nx, d.ny
    om_np = np.random.default_rng(31).uniform(0.4, 1.6, (nx, ny))
    om_o = torch.tensor(om_np[..., None], dtype=torch.float64, requires_grad=True)
    o = P.darcy_c2("oracle", n=14)
    o.omega = om_o
    f = o.assign_fields_sharded() if hasattr(o, "assign_fields_sharded") else None
    f0_np = f.detach().numpy().copy()
    target = torch.randn((nx, ny, 9), generator=torch.Generator().manual_seed(32), dtype=torch.float64)
    for k in range(n_steps):
        f, _ = o.step(f, k)
    J = (0.5 * (f - target) ** 2).sum()
    J.backward()
    om = _cuda(om_np).requires_grad_(True)
    f0 = _cuda(f0_np).permute(2, 0, 1).contiguous()
    fN = darcy_solve(d, om, f0, n_steps, checkpoint_every=K)
    Jg = (0.5 * (fN - target.permute(2, 0, 1).cuda()) ** 2).sum()
    Jg.backward()
    rel = lambda a, b: float((a.cpu().reshape(b.shape) - b).abs().max() / b.abs().max())
    assert rel(fN.permute(1, 2, 0), f.detach()) < 1e-12
    assert rel(om.grad, om_o.grad[..., 0]) < 1e-10, rel(om.grad, om_o.grad[..., 0])


def test_elasticity_gradient_central_differences_16_directions():
    from src.adjoint import elasticity_solve
    nx, n, K = 24, 200, 50
    say("elasticity gradient vs central differences, 16 directions")
    el = P.consolidation_elasticity("gpu", nx=nx)
    ny = el.ny
    lam0, mu0 = (_cuda(t) for t in _fields(nx, ny, 41))
    f0 = torch.zeros((8, nx, ny), dtype=torch.float64, device="cuda")
    target = _rand((8, nx, ny), 42, 1e-4)

    def J(lam, mu):
        fN = elasticity_solve(el, lam, mu, f0, n, checkpoint_every=K)
        return 0.5 * ((fN - target) ** 2).sum()
    mu = mu0.clone().requires_grad_(True)
    J(lam0, mu).backward()
    g = mu.grad.clone()
    eps = 1e-5
    for k in range(16):
        dmu = _rand((nx, ny), 100 + k)
        with torch.no_grad():
            fd = (J(lam0, mu0 + eps * dmu) - J(lam0, mu0 - eps * dmu)).item() / (2 * eps)
        ad = _dot(g, dmu)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad)


def test_steady_state_adjoint_matches_differences_of_converged_solves():
    from src.adjoint import ElasticityStepper, elasticity_steady_adjoint
    nx = 12
    say("fixed-point adjoint of the steady state")
    el = P.consolidation_elasticity("gpu", nx=nx)
    ny = el.ny
    lam0, mu0 = (_cuda(t).reshape(1, nx, ny) for t in _fields(nx, ny, 51))
    target = _rand((8, nx, ny), 52, 1e-4)

    def steady(lam, mu, tol=1e-14):
        st = ElasticityStepper(el, lam, mu)
        f = torch.zeros((8, nx, ny), dtype=torch.float64, device="cuda")
        g = torch.empty_like(f)
        r0 = None
        for k in range(200000):
            st.step(f, g)
            if k % 100 == 0:
                r = float((g - f).norm())
                r0 = r if r0 is None else r0
                if r <= tol * r0:
                    return g
            f, g = g, f
        raise AssertionError("steady state not reached")
    fs = steady(lam0, mu0)
    grads, _, steps, rel = elasticity_steady_adjoint(el, lam0, mu0, fs, fs - target, tol=1e-14)
    say("adjoint steps", steps, "rel", rel)
    eps = 1e-5
    for k in range(3):
        dmu = _rand((1, nx, ny), 60 + k)
        Jp = 0.5 * float(((steady(lam0, mu0 + eps * dmu) - target) ** 2).sum())
        Jm = 0.5 * float(((steady(lam0, mu0 - eps * dmu) - target) ** 2).sum())
        fd = (Jp - Jm) / (2 * eps)
        ad = _dot(grads["mu"], dmu)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad)


def test_reused_solve_buffers_guard_and_private_solves():
    from src.adjoint import elasticity_solve
    say("adjoint session reuse guard")
    nx, n, K = 12, 20, 10
    el = P.consolidation_elasticity("gpu", nx=nx)
    ny = el.ny
    lam, mu1 = (_cuda(t) for t in _fields(nx, ny, 71))
    mu2 = mu1 * 1.1
    f0 = torch.zeros((8, nx, ny), dtype=torch.float64, device="cuda")
    target = _rand((8, nx, ny), 72, 1e-4)

    def J(mu, reuse=True):
        return 0.5 * ((elasticity_solve(el, lam, mu, f0, n, checkpoint_every=K, reuse=reuse) - target) ** 2).sum()
    grads = []
    for m in (mu1, mu2):
        mm = m.clone().requires_grad_(True)
        J(mm).backward()
        grads.append(mm.grad.clone())
    a = mu1.clone().requires_grad_(True)
    b = mu2.clone().requires_grad_(True)
    Ja, Jb = J(a), J(b)
    with pytest.raises(RuntimeError, match="overwritten"):
        Ja.backward()
    a.grad = None
    Ja, Jb = J(a, reuse=False), J(b, reuse=False)
    (Ja + Jb).backward()
    assert torch.equal(a.grad, grads[0]) and torch.equal(b.grad, grads[1])
