"""The oracle against the convergence orders of the reference's own example
drivers (``tests/orders.py``), on the CPU at grid sizes it runs in seconds.  The
orders are the ones the drivers plot; a dropped term or a wrong sign in the
oracle breaks them (the last test plants one)."""
import pytest

import orders as O


def _orders(study, ns, **kw):
    res = [study("oracle", n, **kw) for n in ns]
    return {k: O.order(ns, [r[k] for r in res]) for k in res[0]}, res


def test_darcy_dirichlet_converges_at_first_order():
    # examples/darcy/dirichlet.py:130-141 (the exact solution's time scale uses the
    # padded grid nx, the source the interior n: an O(1/n) term)
    o, res = _orders(O.darcy_dirichlet, (16, 32))
    assert abs(o["l2"] - 1.0) < 0.2 and abs(o["linf"] - 1.0) < 0.2, (o, res)


@pytest.mark.parametrize("walls,eta_order", [("dirichlet", 2.0), ("neumann", 1.0)])
def test_elasticity_walls_converge(walls, eta_order):
    # examples/elasticity/{dirichlet,neumann}.py:169-182: displacement second order
    # with full-way Dirichlet walls, first with traction walls; stress first order
    o, res = _orders(O.elasticity_walls, (16, 32), walls=walls)
    assert abs(o["eta"] - eta_order) < 0.2 and abs(o["sigma"] - 1.0) < 0.2, (o, res)


def test_biot_manufactured_solution_converges_at_second_order():
    # examples/poroelasticity/periodic.py:297-307: O(dx^2) for displacement, stress, pressure
    o, res = _orders(O.biot_periodic, (12, 16))
    assert all(abs(v - 2.0) < 0.2 for v in o.values()), (o, res)
    assert res[-1]["eta"] < 0.06 and res[-1]["p"] < 0.03, res


def test_terzaghi_converges_at_first_order():
    # examples/poroelasticity/Terzaghi.py:378-386: O(dx) for subsidence and pressure
    # with the driver's default single grid (the multigrid form of the paper's sweep is
    # still pre-asymptotic at 8 -> 16, slope 1.23; the GPU tests fit it on 16 ... 64)
    o, res = _orders(O.terzaghi, (8, 16))
    assert abs(o["subsidence"] - 1.0) < 0.2 and abs(o["p"] - 1.0) < 0.2, (o, res)


def test_biot_order_detects_a_wrong_gradient_sign(monkeypatch):
    # a mistake planted in the oracle (the D2Q9 lattice gradient with the wrong
    # sign, as a pull from x + c instead of x - c would give) breaks the pin above
    from oracle import darcy as OD
    grad = OD.OracleDarcy.gradient
    monkeypatch.setattr(OD.OracleDarcy, "gradient", lambda self, q, scale=1.0: grad(self, q, -scale))
    o, res = _orders(O.biot_periodic, (12, 16))
    second_order = all(abs(v - 2.0) < 0.2 for v in o.values()) and res[-1]["eta"] < 0.06
    assert not second_order, (o, res)
