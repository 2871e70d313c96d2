"""Step loops to a relative step-to-step residual with the test on the device
(``run_steps_to``): the step count, the latched state and the ratio against the
same loop written on the host (through ``run_steps``) and against the oracle's
loop; two calls in a row replay the captured blocks with eager work between."""
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def host_loop(run1, f0, tol, ce, norm):
    f1 = run1(f0, 1)
    r0 = norm(f1, f0)
    f, k = f1, 1
    while True:
        f = run1(f, ce - 1)
        fn = run1(f, 1)
        r = norm(fn, f)
        f, k = fn, k + ce
        if r / r0 < tol:
            return f, k, r / r0


def gnorm(a, b):
    return torch.linalg.norm((a - b).double()).item()


@pytest.mark.parametrize("tol,ce,checks", [(1e-6, 10, 4), (1e-8, 6, 3)])
def test_elasticity_steps_to_tolerance(tol, ce, checks):
    say("elasticity run_steps_to", tol, ce, checks)
    g = P.periodic_elasticity("gpu", nx=32)
    f0 = g.assign_fields_sharded()
    fh, kh, rh = host_loop(lambda f, n: g.run_steps(f, n), f0, tol, ce, gnorm)
    for rep in range(2):
        fd, kd, rd = g.run_steps_to(f0, tol, check_every=ce, checks=checks)
        say("  host", kh, rh, "device", kd, rd)
        assert kd == kh
        assert torch.equal(fd, fh)
        assert abs(rd - rh) <= 1e-10 * rh
        g.run_steps(f0, 3)  # eager work between the replays
    # the oracle's loop (elasticity.py:1169-1219 steps) reaches the same count and state
    o = P.periodic_elasticity("oracle", nx=32)

    def orun(f, n):
        for _ in range(n):
            f, _ = o._step(f, 0)
        return f
    fo, ko, ro = host_loop(orun, o.assign_fields_sharded(), tol, ce, gnorm)
    assert ko == kd, (ko, kd)
    assert (fd.cpu() - fo).abs().max().item() <= 1e-12 * fo.abs().max().item()


def test_elasticity_steps_to_cap():
    g = P.periodic_elasticity("gpu", nx=16)
    f0 = g.assign_fields_sharded()
    fd, kd, rd = g.run_steps_to(f0, 1e-30, check_every=10, max_steps=100, checks=2)
    assert kd == -1 and rd > 0


def test_darcy_steps_to_tolerance():
    say("darcy run_steps_to")
    d = P.darcy_c2("gpu", n=32)
    f0 = d.assign_fields_sharded()
    fh, kh, rh = host_loop(lambda f, n: d.run_steps(f, n), f0, 1e-8, 10, gnorm)
    for rep in range(2):
        fd, kd, rd = d.run_steps_to(f0, 1e-8, check_every=10, checks=5)
        say("  host", kh, rh, "device", kd, rd)
        assert kd == kh
        assert torch.equal(fd, fh)
        assert abs(rd - rh) <= 1e-10 * rh
        d.run_steps(f0, 5)
