"""Pins of the CPU oracle (no GPU): closed forms, invariants, brute force on tiny
grids and convergence orders, each chosen so that a dropped term, a wrong sign or
a transposed index fails."""
import numpy as np
import pytest
import torch

import problems as P
from oracle import bc as obc
from oracle.darcy import OracleDarcy
from oracle.elasticity import OracleElasticity
from oracle.lattice import D2Q8, D2Q9

torch.set_num_threads(4)


# ---- lattices (reference lattice.py:124-168) -------------------------------------
def test_d2q8_tables():
    L = D2Q8()
    assert L.c[0].tolist() == [1, 0, -1, 0, 1, -1, -1, 1]
    assert L.c[1].tolist() == [0, 1, 0, -1, 1, 1, -1, -1]
    assert L.opp_indices.tolist() == [2, 3, 0, 1, 6, 7, 4, 5]
    assert L.right_indices.tolist() == [0, 4, 7] and L.left_indices.tolist() == [2, 5, 6]
    assert L.main_indices.tolist() == [0, 1, 2, 3]
    np.testing.assert_allclose(L.w, [1 / 9] * 4 + [1 / 36] * 4)
    np.testing.assert_allclose(L.w @ L.c.T, 0.0, atol=1e-16)


def test_d2q9_tables_isotropy():
    L = D2Q9()
    assert L.opp_indices.tolist() == [0, 2, 1, 6, 5, 4, 3, 8, 7]
    np.testing.assert_allclose(L.w.sum(), 1.0)
    np.testing.assert_allclose(L.w @ L.c.T, 0.0, atol=1e-16)
    M = 3 * np.einsum("i,ai,bi->ab", L.w, L.c, L.c)
    np.testing.assert_allclose(M, np.eye(2), atol=1e-15)


def _el(lam=1.0, mu=1.0, nx=4, ny=4, **kw):
    return OracleElasticity(**{"lambda": lam, "mu": mu, "nx": nx, "ny": ny, "nz": 0, "precision": "f64/f64",
                               "lattice": D2Q8(), **kw})


# ---- transfer matrices and pointwise maps (elasticity.py:805-1058) ---------------
@pytest.mark.parametrize("lam,mu", [(1.0, 1.0), (2.0, 1.0), (0.2444, 0.0306)])
def test_transfer_matrix_identity(lam, mu):
    el = _el(lam, mu)
    prod = el.m_to_f @ el.f_to_m  # (7, 8) x (8, 5)
    np.testing.assert_allclose(prod[:5].numpy(), np.eye(5), atol=1e-15)
    np.testing.assert_allclose(prod[5:].numpy(), 0.0, atol=1e-15)


def test_collision_relaxes_moments():
    """f_post @ f_to_m = (m10+F, m01+F, cs m_s, cd m_d, cd m_11) and the third-order
    moments sum c_x c_y^2 f = m10/3 + F/6, sum c_x^2 c_y f = m01/3 + F/6."""
    lam, mu = 1.7, 0.8
    el = _el(lam, mu, nx=3, ny=2)
    rng = np.random.default_rng(0)
    f = torch.as_tensor(rng.standard_normal((3, 2, 8)))
    F = torch.as_tensor(rng.standard_normal((3, 2, 2)))
    fp = el.collision(f, F, use_force=True)
    m = f @ el.f_to_m
    mp = fp @ el.f_to_m
    cs = (3 * (lam + mu) - 1) / (3 * (lam + mu) + 1)
    cd = (6 * mu - 1) / (6 * mu + 1)
    np.testing.assert_allclose(mp[..., 0:2], m[..., 0:2] + F, atol=1e-14)
    np.testing.assert_allclose(mp[..., 2], cs * m[..., 2], atol=1e-14)
    np.testing.assert_allclose(mp[..., 3:5], cd * m[..., 3:5], atol=1e-14)
    cx = torch.tensor([1, 0, -1, 0, 1, -1, -1, 1], dtype=torch.float64)
    cy = torch.tensor([0, 1, 0, -1, 1, 1, -1, -1], dtype=torch.float64)
    np.testing.assert_allclose(fp @ (cx * cy ** 2), m[..., 0] / 3 + F[..., 0] / 6, atol=1e-14)
    np.testing.assert_allclose(fp @ (cx ** 2 * cy), m[..., 1] / 3 + F[..., 1] / 6, atol=1e-14)


def test_equilibrium_macroscopic_round_trip():
    el = _el(1.3, 0.4, nx=5, ny=3)
    rng = np.random.default_rng(1)
    eta = torch.as_tensor(rng.standard_normal((5, 3, 2)))
    sig = torch.as_tensor(rng.standard_normal((5, 3, 3)))
    F = torch.as_tensor(rng.standard_normal((5, 3, 2)))
    f = el.equilibrium(eta, sig, external_force=F)
    e2, s2 = el.update_macroscopic(f, F)
    np.testing.assert_allclose(e2, eta, atol=1e-14)
    np.testing.assert_allclose(s2, sig, atol=1e-14)
    # divergence is the trace part: div = -(sxx+syy) / (2(lam+mu)) ... via m_s
    div = el.update_divergence(f)[..., 0]
    m_s = f @ el.f_to_m[:, 2]
    np.testing.assert_allclose(div, -1.5 / (3 * 1.7 + 1) * m_s, atol=1e-14)


def test_streaming_is_numpy_roll():
    L = D2Q9()
    rng = np.random.default_rng(2)
    f = rng.standard_normal((3, 4, 9))
    out = obc.stream(torch.as_tensor(f), L).numpy()
    for i in range(9):
        for x in range(3):
            for y in range(4):
                assert out[x, y, i] == f[(x - L.c[0, i]) % 3, (y - L.c[1, i]) % 4, i]
    np.testing.assert_allclose(out.sum(axis=(0, 1)), f.sum(axis=(0, 1)))


def test_restrict_prolong():
    el = _el(nx=8, ny=8, MG_levels=2)
    rng = np.random.default_rng(3)
    c = torch.as_tensor(rng.standard_normal((4, 4, 8)))
    np.testing.assert_allclose(el._restrict(el._prolong(c)), 4 * c, atol=1e-14)
    f = torch.as_tensor(rng.standard_normal((8, 8, 8)))
    r = el._restrict(f)
    for X in range(4):
        for Y in range(4):
            np.testing.assert_allclose(r[X, Y], f[2 * X:2 * X + 2, 2 * Y:2 * Y + 2].sum(axis=(0, 1)), atol=1e-14)
    p = el._prolong(c)
    assert torch.equal(p[5, 2], c[2, 1])


def test_c1_second_order_convergence():
    errs = []
    for nx in (16, 32):
        el = P.periodic_elasticity("oracle", nx=nx)
        f = el.assign_fields_sharded()
        r0 = None
        for k in range(5000):
            fn, _ = el._step(f, k)
            r = torch.linalg.norm(fn - f).item()
            r0 = r if r0 is None else r0
            f = fn
            if r / r0 < 1e-10:
                break
        eta, _ = el.update_macroscopic(f)
        ue = P.manufactured_c1(nx) / (8 * np.pi ** 2)
        errs.append(np.linalg.norm(eta.numpy() - ue) / np.linalg.norm(ue))
    order = np.log2(errs[0] / errs[1])
    assert 1.9 < order < 2.1, (errs, order)


def test_vcycle_fixed_point_and_residual():
    """A converged state is a fixed point of the damped smoother and of the V-cycle,
    and the V-cycle's norm is ||f - S(f)|| of its pre-smoothed state."""
    el1 = P.periodic_elasticity("oracle", nx=16)
    f = el1.assign_fields_sharded()
    for k in range(3000):
        f, _ = el1._step(f, k)
    elm = P.periodic_elasticity("oracle", nx=16, levels=3)
    fm, res = elm.step(f, 0)
    assert res.item() < 1e-13
    np.testing.assert_allclose(fm, f, atol=1e-14)
    # norm reported = |f_s - S(f_s)| where f_s is f after k pre-smoothing steps
    g = torch.zeros_like(f)
    gs = g
    for _ in range(elm.k_smooth):
        gs = (1 - elm.gamma) * gs + elm.gamma * elm._step(gs, 0)[0]
    _, res = elm.step(g, 0)
    np.testing.assert_allclose(res.item(), torch.linalg.norm(gs - elm._step(gs, 0)[0]).item(), rtol=1e-13)


def test_vcycle_converges_to_plain_solution():
    """V-cycles reach the plain solver's fixed point (same discretisation error to the
    closed form).  The reference V-cycle has no exact coarse solve: with a 16^2
    coarsest grid it contracts, with an 8^2 one and k=2 it stagnates."""
    el = P.periodic_elasticity("oracle", nx=64, levels=3)
    f = el.assign_fields_sharded()
    res = []
    for t in range(30):
        f, r = el.step(f, t)
        res.append(r.item())
    assert res[-1] / res[0] < 1e-8
    eta, _ = el.update_macroscopic(f)
    ue = P.manufactured_c1(64) / (8 * np.pi ** 2)
    err = np.linalg.norm(eta.numpy() - ue) / np.linalg.norm(ue)
    assert abs(err - 0.0016060) < 2e-5, err


@pytest.mark.parametrize("levels", [2, 3])
def test_vcycle_with_coarse_obstacles_reaches_the_single_grid_fixed_point(levels):
    """Interior solid obstacles on every multigrid level (each level's own wall and
    obstacle BCs, elasticity.py:396-436): the V-cycles converge, and to a fixed point
    of the single-grid step with the fine obstacle (S(f) = f on the interior box the
    residual covers; the wall layers are the walls' own), the discretisation the
    coarse levels only accelerate.  (The plain step alone keeps the full-way walls'
    period-2 mode, so the fixed point is reached through the damped V-cycle.)"""
    om = P.obstacle_elasticity("oracle", levels=levels)
    o1 = P.obstacle_elasticity("oracle")
    f = om.assign_fields_sharded()
    x0, x1, y0, y1 = 1, om.nx - 1, 1, om.ny - 1
    inner = lambda g: g[x0:x1, y0:y1]
    s0 = torch.linalg.norm(inner(o1._step(torch.zeros_like(f), 0)[0])).item()
    res = []
    for t in range(120):
        f, r = om.step(f, t)
        res.append(r.item())
    assert res[-1] < 1e-11 * res[0], res[::10]
    assert all(b < a for a, b in zip(res[::10], res[10::10])), res[::10]
    assert torch.linalg.norm(inner(o1._step(f, 0)[0] - f)).item() < 1e-10 * s0


# ---- Darcy (darcy.py) ---------------------------------------------------------------
def _darcy(nx=16, ny=16, omega=1.0 / (3 * 0.4 + 0.5)):
    return OracleDarcy(omega=omega, nx=nx, ny=ny, nz=0, precision="f64/f64", lattice=D2Q9())


def test_darcy_mass_conservation_and_rest_state():
    d = _darcy()
    rng = np.random.default_rng(4)
    f = torch.as_tensor(rng.random((16, 16, 9)))
    m0 = f.sum().item()
    for t in range(10):
        f, _ = d.step(f, t)
    np.testing.assert_allclose(f.sum().item(), m0, rtol=1e-13)
    rest = torch.as_tensor(np.broadcast_to(0.7 * D2Q9().w, (16, 16, 9)).copy())
    np.testing.assert_allclose(d.step(rest, 0)[0], rest, atol=1e-16)


def test_darcy_mode_decay():
    """A periodic cosine mode decays at the rate K k^2 of the diffusion equation,
    K = (1/omega - 1/2)/3; the rate error is second order in k (halving k quarters it)."""
    K = 0.4
    errs = []
    for N, T in ((64, 400), (128, 1600)):
        d = _darcy(N, 4, 1.0 / (3 * K + 0.5))
        k = 2 * np.pi / N
        x = np.arange(N)[:, None, None]
        f = torch.as_tensor((1 + 1e-2 * np.cos(k * x) * np.ones((1, 4, 1))) * D2Q9().w)
        for t in range(T):
            f, _ = d.step(f, t)
        p = d.update_macroscopic(f)[..., 0].numpy()
        amp = (p[:, 0] - 1) @ np.cos(k * np.arange(N)) * 2 / N
        rate = -np.log(amp / 1e-2) / T
        errs.append(abs(rate / (K * k ** 2) - 1))
    assert errs[0] < 0.01 and 3.5 < errs[0] / errs[1] < 4.5, errs


def test_gradient_spectral_symbol():
    """3 sum_i w_i c_i q(x + c_i) of q = sin(kx) is sin(k) cos(kx), exactly."""
    N = 24
    d = _darcy(N, 6)
    k = 2 * np.pi / N
    x = np.arange(N)[:, None, None] + 0 * np.arange(6)[None, :, None]
    q = torch.as_tensor(np.sin(k * x))
    g = d.gradient(q, 0.7).numpy()
    np.testing.assert_allclose(g[..., 0], 0.7 * np.sin(k) * np.cos(k * x[..., 0]), atol=1e-15)
    np.testing.assert_allclose(g[..., 1], 0.0, atol=1e-15)


def test_grf_statistics():
    K = P.lognormal_field(256, 256, 1.0, 16.0, 0)
    g = np.log(K)
    assert abs(g.mean()) < 0.3 and 0.6 < g.var() < 1.4
    # correlation at lag 16 ~ exp(-1)
    c16 = np.mean((g[:-16] - g.mean()) * (g[16:] - g.mean())) / g.var()
    assert 0.2 < c16 < 0.55
