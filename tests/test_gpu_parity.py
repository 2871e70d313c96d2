"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same inputs.

Tolerances (relative, max-norm over the compared array): 1e-12 in FP64, 1e-5 in FP32.
"""
import numpy as np
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu

TOL = {"f64/f64": 1e-12, "f32/f32": 1e-5}


def rel(a, b, floor=0.0):
    """max|a-b| / max(max|b|, floor): ``floor`` is the scale below which a field is
    round-off (fields that are identically zero in exact arithmetic)."""
    a = torch.as_tensor(a).detach().to("cpu", torch.float64)
    b = torch.as_tensor(b).detach().to("cpu", torch.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    den = max(b.abs().max().item(), floor)
    return (a - b).abs().max().item() / (den if den > 0 else 1.0)


def say(*args):
    print(*args, flush=True)


@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
def test_elasticity_periodic_steps(precision):
    say("elasticity periodic", precision)
    g = P.periodic_elasticity("gpu", nx=64, precision=precision)
    o = P.periodic_elasticity("oracle", nx=64, precision=precision)
    fg, fo = g.assign_fields_sharded(), o.assign_fields_sharded()
    for t in range(40):
        fg, pg = g.step(fg, t, return_fpost=(t == 39))
        fo, po = o.step(fo, t, return_fpost=(t == 39))
    assert rel(fg, fo) < TOL[precision]
    assert rel(pg, po) < TOL[precision]
    eg, sg = g.update_macroscopic(fg)
    eo, so = o.update_macroscopic(fo)
    assert rel(eg, eo) < TOL[precision]
    assert rel(sg, so) < TOL[precision]


def test_elasticity_periodic_nonsquare_and_ragged():
    # ny not a multiple of the 62-cell strip, odd nx: partial strips and row blocks
    mg, mo = P.modules("gpu"), P.modules("oracle")
    for nx, ny in ((37, 75), (5, 130), (1, 3), (64, 62), (3, 63)):
        say("elasticity ragged", nx, ny)
        kw = {"lambda": 1.3, "mu": 0.6, "nx": nx, "ny": ny, "nz": 0, "precision": "f64/f64"}
        g = mg["Elasticity"](lattice=mg["D2Q8"]("f64/f64"), print_info_rate=0, **kw)
        o = mo["Elasticity"](lattice=mo["D2Q8"](), **kw)
        rng = np.random.default_rng(7)
        f0 = rng.standard_normal((nx, ny, 8)) * 1e-3
        fg = torch.as_tensor(f0, device="cuda")
        fo = torch.as_tensor(f0)
        ext = rng.standard_normal((nx, ny, 2)) * 1e-4
        for t in range(7):
            fg, _ = g.step(fg, t, external_force=torch.as_tensor(ext, device="cuda"))
            fo, _ = o.step(fo, t, external_force=torch.as_tensor(ext))
        assert rel(fg, fo) < 1e-12, (nx, ny)


@pytest.mark.parametrize("threads", [0, 256, 704])
def test_elasticity_interior_obstacle(threads):
    """An interior L-shaped solid obstacle with a prescribed displacement (halfway
    bounce-back on solid cells) inside full-way Dirichlet walls: single steps, then
    tile launches of 4 steps, against the oracle's literal boundary passes."""
    say("elasticity interior obstacle, tile threads", threads)
    g, o = P.obstacle_elasticity("gpu"), P.obstacle_elasticity("oracle")
    rng = np.random.default_rng(9)
    f0 = rng.standard_normal((g.nx, g.ny, 8)) * 1e-3
    fg = torch.as_tensor(f0, device="cuda")
    fo = torch.as_tensor(f0)
    for t in range(12):
        fg, _ = g.step(fg, t)
        fo, _ = o.step(fo, t)
    assert rel(fg, fo) < 1e-12
    fs = g._soa(fg, 8)
    for _ in range(3):
        fs = g._stepn_into(fs, torch.empty_like(fs), 4, threads=threads)
    for t in range(12, 24):
        fo, _ = o.step(fo, t)
    assert rel(fs.permute(1, 2, 0), fo) < 1e-12


@pytest.mark.parametrize("name,precision", [("neumann", "f64/f64"), ("dirichlet", "f64/f64"),
                                            ("mixed", "f64/f64"), ("neumann", "f32/f32"), ("mixed", "f32/f32")])
def test_elasticity_walls_steps(name, precision):
    say("elasticity walls", name, precision)
    make = {"neumann": P.neumann_elasticity, "dirichlet": P.dirichlet_elasticity,
            "mixed": P.mixed_halfway_elasticity}[name]
    g, o = make("gpu", precision=precision), make("oracle", precision=precision)
    rng = np.random.default_rng(3)
    f0 = rng.standard_normal((g.nx, g.ny, 8)) * 1e-3
    fg = torch.as_tensor(f0, device="cuda").to(g.precisionPolicy.torch_compute)
    fo = torch.as_tensor(f0).to(o.precisionPolicy.compute_dtype)
    for t in range(25):
        fg, _ = g.step(fg, t)
        fo, _ = o.step(fo, t)
    assert rel(fg, fo) < TOL[precision]


@pytest.mark.parametrize("name,levels", [("periodic", 3), ("neumann", 2), ("neumann", 3), ("dirichlet", 3),
                                         ("obstacle", 2), ("obstacle", 3)])
def test_elasticity_vcycle(name, levels):
    # "obstacle": an interior L-shaped obstacle on every level (BCs_MG with coarse voxels)
    say("elasticity V-cycle", name, levels)
    if name == "periodic":
        g, o = P.periodic_elasticity("gpu", nx=32, levels=levels), P.periodic_elasticity("oracle", nx=32, levels=levels)
    else:
        make = {"neumann": P.neumann_elasticity, "dirichlet": P.dirichlet_elasticity,
                "obstacle": P.obstacle_elasticity}[name]
        g, o = make("gpu", n=16, levels=levels), make("oracle", n=16, levels=levels)
    fg, fo = g.assign_fields_sharded(), o.assign_fields_sharded()
    for t in range(4):
        fg, rg = g.step(fg, t)
        fo, ro = o.step(fo, t)
        assert abs(rg.item() - ro.item()) <= 1e-12 * max(abs(ro.item()), 1e-300), (t, rg.item(), ro.item())
    assert rel(fg, fo) < 1e-12


def test_elasticity_macro_div_equilibrium():
    say("elasticity macro/div/equilibrium")
    g = P.neumann_elasticity("gpu", n=16)
    o = P.neumann_elasticity("oracle", n=16)
    rng = np.random.default_rng(11)
    eta = rng.standard_normal((g.nx, g.ny, 2))
    sig = rng.standard_normal((g.nx, g.ny, 3))
    ext = rng.standard_normal((g.nx, g.ny, 2)) * 1e-2
    fg = g.equilibrium(torch.as_tensor(eta, device="cuda"), torch.as_tensor(sig, device="cuda"),
                       external_force=torch.as_tensor(ext, device="cuda"))
    fo = o.equilibrium(torch.as_tensor(eta), torch.as_tensor(sig), external_force=torch.as_tensor(ext))
    assert rel(fg, fo) < 1e-13
    eg, sg = g.update_macroscopic(fg, torch.as_tensor(ext, device="cuda"))
    eo, so = o.update_macroscopic(fo, torch.as_tensor(ext))
    assert rel(eg, eo) < 1e-13 and rel(sg, so) < 1e-13
    # round trip of the inverse pair
    assert rel(eg, torch.as_tensor(eta)) < 1e-12 and rel(sg, torch.as_tensor(sig)) < 1e-12
    assert rel(g.update_divergence(fg), o.update_divergence(fo)) < 1e-13
    cg = g.collision(fg, torch.as_tensor(ext, device="cuda"))
    co = o.collision(fo, torch.as_tensor(ext))
    assert rel(cg, co) < 1e-13


@pytest.mark.parametrize("precision,hetero,source", [("f64/f64", True, False), ("f32/f32", True, False),
                                                      ("f64/f64", False, True)])
def test_darcy_steps(precision, hetero, source):
    say("darcy C2-like", precision, hetero, source)
    g = P.darcy_c2("gpu", n=40, precision=precision, heterogeneous=hetero, source=source)
    o = P.darcy_c2("oracle", n=40, precision=precision, heterogeneous=hetero, source=source)
    fg, fo = g.assign_fields_sharded(), o.assign_fields_sharded()
    assert rel(fg, fo) < TOL[precision]
    ext = 1e-3 * np.sin(np.arange(g.nx * g.ny).reshape(g.nx, g.ny, 1))
    for t in range(30):
        e = ext if t % 2 else None
        fg, pg = g.step(fg, t, return_fpost=(t == 29),
                        external_source=None if e is None else torch.as_tensor(e, device="cuda"))
        fo, po = o.step(fo, t, return_fpost=(t == 29),
                        external_source=None if e is None else torch.as_tensor(e).to(o.precisionPolicy.compute_dtype))
    assert rel(fg, fo) < TOL[precision]
    assert rel(pg, po) < TOL[precision]
    rg, gg = g.update_macroscopic(fg, gradient=True)
    ro, go = o.update_macroscopic(fo, gradient=True)
    assert rel(rg, ro) < TOL[precision] and rel(gg, go) < TOL[precision]
    for scale in (1.0, 0.37):
        assert rel(g.gradient(rg, scale), o.gradient(ro, scale)) < TOL[precision]


@pytest.mark.parametrize("threads", [0, 256, 704])
def test_darcy_interior_obstacle(threads):
    """An arbitrary solid voxel set (an L-shaped obstacle with a Dirichlet surface)
    inside the C2 walls: single steps, then tile launches of 4 steps, against the
    oracle's literal boundary passes."""
    say("darcy interior obstacle, tile threads", threads)
    g, o = P.darcy_obstacle("gpu"), P.darcy_obstacle("oracle")
    fg, fo = g.assign_fields_sharded(), o.assign_fields_sharded()
    for t in range(12):
        fg, _ = g.step(fg, t)
        fo, _ = o.step(fo, t)
    assert rel(fg, fo) < 1e-12
    fs = g._soa(fg, 9)
    for _ in range(3):
        fs = g._stepn_into(fs, torch.empty_like(fs), 4, threads=threads)
    for t in range(12, 24):
        fo, _ = o.step(fo, t)
    assert rel(fs.permute(1, 2, 0), fo) < 1e-12


def test_darcy_equilibrium_walls():
    say("darcy equilibrium walls")
    g, o = P.darcy_equilibrium("gpu"), P.darcy_equilibrium("oracle")
    fg, fo = g.assign_fields_sharded(), o.assign_fields_sharded()
    assert rel(fg, fo) < 1e-14
    for t in range(20):
        fg, _ = g.step(fg, t)
        fo, _ = o.step(fo, t)
    assert rel(fg, fo) < 1e-12


@pytest.mark.parametrize("levels,coupling,alt,fused", [(1, "centered", False, True), (1, "implicit", True, True),
                                                       (3, "centered", False, True), (2, "explicit", False, True),
                                                       (1, "centered", False, False), (3, "centered", False, False)])
def test_biot_periodic_steps(levels, coupling, alt, fused):
    say("biot periodic", levels, coupling, alt, fused)
    g = P.biot_periodic("gpu", nx=32, levels=levels, steps=2, coupling=coupling, alt=alt)
    g.fusedCoupling = fused
    o = P.biot_periodic("oracle", nx=32, levels=levels, steps=2, coupling=coupling, alt=alt)
    sg = g.initial_state()
    so = o.initial_state()
    for a, b in zip(sg, so):
        assert rel(a.permute(1, 2, 0), b, 1e-3) < 1e-12
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(3):
        outg = g.step(fe, fd, dde, t)
        outo = o.step(oe, od, odde, t)
        errs = P.step_rel(outg, outo)
        assert max(errs) < 1e-12, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]


def test_biot_consolidation_mg_walls_fp32():
    # the consolidation set-up in FP32 state (the precision of the gradient
    # workload) against the FP32 oracle: 1e-5; the divergence rate is measured
    # against the populations' scale (it is formed from a difference of states)
    say("biot consolidation MG walls, FP32")
    g = P.biot_consolidation("gpu", nx=32, levels=3, steps=2, precision="f32/f32")
    o = P.biot_consolidation("oracle", nx=32, levels=3, steps=2, precision="f32/f32")
    sg = g.initial_state()
    so = o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(3):
        outg = g.step(fe, fd, dde, t)
        outo = o.step(oe, od, odde, t)
        errs = P.step_rel(outg, outo, against_state=(4,))
        say("  step", t, "relative errors", [f"{e:.1e}" for e in errs])
        assert max(errs) < 1e-5, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]


@pytest.mark.parametrize("fused", [True, False])
def test_biot_consolidation_mg_walls(fused):
    say("biot consolidation MG walls, fused coupling force:", fused)
    g = P.biot_consolidation("gpu", nx=32, levels=3, steps=2)
    g.fusedCoupling = fused
    o = P.biot_consolidation("oracle", nx=32, levels=3, steps=2)
    sg = g.initial_state()
    so = o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(3):
        outg = g.step(fe, fd, dde, t)
        outo = o.step(oe, od, odde, t)
        errs = P.step_rel(outg, outo)
        assert max(errs) < 1e-12, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]


def test_c1_converges_to_closed_form():
    say("C1 64^2 to residual 1e-10")
    nx = 64
    g = P.periodic_elasticity("gpu", nx=nx)
    f = g.assign_fields_sharded()
    r0 = None
    for k in range(20000):
        fn, _ = g.step(f, k)
        if k % 10 == 0:
            r = torch.linalg.norm(fn - f).item()
            r0 = r if r0 is None else r0
            if r / r0 < 1e-10:
                f = fn
                break
        f = fn
    eta, _ = g.update_macroscopic(f)
    ue = P.manufactured_c1(nx) / (8 * np.pi ** 2)
    err = np.linalg.norm(eta.cpu().numpy() - ue) / np.linalg.norm(ue)
    # second order: 0.0256 / 0.0064 / 0.0016 at 16 / 32 / 64 cells (oracle)
    assert 1.4e-3 < err < 1.8e-3, err


def test_vcycle_graph_replays_match_eager():
    say("V-cycle graph vs eager")
    g = P.neumann_elasticity("gpu", n=32, levels=3)
    f = g.assign_fields_sharded()
    vg = g.vcycle_graph()
    vg.load(g._soa(f, 8))
    for t in range(4):
        n_graph = vg.replay().item()
        f, n_eager = g.step(f, t)
        torch.cuda.synchronize()
        assert rel(vg.state.permute(1, 2, 0), f) < 1e-15
        assert abs(n_graph - n_eager.item()) <= 1e-15 * abs(n_eager.item())


def test_biot_residual_driven_inner_loop_matches_oracle():
    say("biot inner loop to tolerance")
    # material of the reference's periodic Biot driver (E = 0.11, nu = 0.8, K = 0.1), for
    # which the V-cycle contracts (at lambda = mu = 1 it does not, see the bench)
    lam, mu = 0.11 * 0.8 / (1 - 0.8 ** 2), 0.5 * 0.11 / 1.8
    g = P.biot_periodic("gpu", nx=64, levels=4, steps=1, lam=lam, mu=mu, K=0.1)
    g.elasticityTol = 1e-3
    g.elasticityMaxSteps = 200
    o = P.biot_periodic("oracle", nx=64, levels=4, steps=1, lam=lam, mu=mu, K=0.1)
    sg, so = g.initial_state(), o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(2):
        outg = g.step(fe, fd, dde, t)
        o.elasticitySteps = g.inner_counts[-1]
        outo = o.step(oe, od, odde, t)
        assert g.inner_counts[-1] > 1
        errs = P.step_rel(outg, outo)
        assert max(errs) < 1e-12, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]


@pytest.mark.parametrize("levels", [1, 3])
def test_run_loop_matches_oracle(levels):
    # run(t_max) performs t_max + 1 steps (elasticity.py:1372); without I/O the loop
    # replays captured graphs
    say("elasticity run()", levels)
    # 64^2 with 3 levels: a contracting V-cycle (coarsest grid 16^2)
    g = P.periodic_elasticity("gpu", nx=64, levels=levels)
    o = P.periodic_elasticity("oracle", nx=64, levels=levels)
    g.ioRate = 0
    g.printInfoRate = 0
    fg = g.run(20)
    fo = o.run(20)
    assert rel(fg, fo) < 1e-12


@pytest.mark.parametrize("io", [0, 2])
def test_poroelasticity_run_matches_oracle(io):
    # LBMPoroelasticity.run (poroelasticity.py:208-301): t_max + 1 physical steps; with
    # io_rate 0 the inner iterations replay captured graphs, with io_rate 2 every second
    # step hands its seven fields to output_data
    say("poroelasticity run(), io rate", io)
    g = P.biot_periodic("gpu", nx=32, levels=2, steps=3)
    o = P.biot_periodic("oracle", nx=32, levels=2, steps=3)
    seen = {"gpu": [], "oracle": []}
    for kind, b in (("gpu", g), ("oracle", o)):
        b.ioRate = io
        # fields arrive on the host with process_allgather's leading process axis
        b.output_data = (lambda kind: lambda **kw: seen[kind].append(
            (kw["timestep"], torch.as_tensor(np.array(kw["pressure"])[0]).double().clone(),
             tuple(np.shape(kw["eta"])), isinstance(kw["eta"], np.ndarray))))(kind)
    ge, gd = g.run(5)
    oe, od = o.run(5)
    assert all(sh == (1, 32, 32, 2) and host for _, _, sh, host in seen["gpu"])
    seen = {k: [x[:2] for x in v] for k, v in seen.items()}
    assert rel(ge, oe) < 1e-12 and rel(gd, od) < 1e-12
    assert [t for t, _ in seen["gpu"]] == [t for t, _ in seen["oracle"]] == ([0, 2, 4, 5] if io else [])
    for (_, a), (_, b) in zip(seen["gpu"], seen["oracle"]):
        assert rel(a, b, 1e-3) < 1e-12


def test_darcy_run_loop_matches_oracle():
    say("darcy run()")
    g = P.darcy_c2("gpu", n=30)
    o = P.darcy_c2("oracle", n=30)
    g.printInfoRate = 0
    assert rel(g.run(150), o.run(150)) < 1e-12


@pytest.mark.parametrize("precision", ["f64/f64", "f32/f32"])
def test_c2_full_size_matches_oracle(precision):
    """C2 at its full size (512 x 512 cells with the ghost walls, log-normal K of
    variance 1 and correlation 16, seed 0) through the bench's launch path: the tuned
    multi-step tile launches replayed from captured graphs (run_steps), against 24
    steps of the oracle."""
    say("C2 514^2 run_steps vs oracle", precision)
    g = P.darcy_c2("gpu", n=512, precision=precision, heterogeneous=True, seed=0, variance=1.0, corr=16.0)
    o = P.darcy_c2("oracle", n=512, precision=precision, heterogeneous=True, seed=0, variance=1.0, corr=16.0)
    g.autotune()
    say("launch", g._multi()[:2])
    fg = g.run_steps(g.assign_fields_sharded(), 24, chunk=12)
    fo = o.assign_fields_sharded()
    for t in range(24):
        fo, _ = o.step(fo, t + 1)
    assert rel(fg, fo) < TOL[precision]


def test_c3_full_size_matches_oracle():
    """C3 at its size and depth: periodic Biot 256 x 256, elasticity V-cycles of 4
    levels with k = 2 and lambda = mu = 1 (as specified), FP64; three physical steps of
    two V-cycles each against the oracle, all seven outputs to 1e-12."""
    say("C3 256^2, 4 levels, k = 2, lambda = mu = 1")
    g = P.biot_periodic("gpu", nx=256, levels=4, steps=2, k_smooth=2)
    o = P.biot_periodic("oracle", nx=256, levels=4, steps=2, k_smooth=2)
    sg = g.initial_state()
    so = o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(3):
        outg = g.step(fe, fd, dde, t)
        outo = o.step(oe, od, odde, t)
        errs = P.step_rel(outg, outo)
        say("  step", t, "relative errors", [f"{e:.1e}" for e in errs])
        assert max(errs) < 1e-12, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]


def test_c4_full_size_forward_matches_oracle_fp32():
    """C4's forward at its size and depth: the 2-D consolidation 1024 x 1026, 10 levels,
    k = 8, FP32 state, lambda = 2 and mu = 1 held as fields, the Darcy rate from the
    seeded log-normal K (variance 2, correlation 32); one physical step of two V-cycles
    against the FP32 oracle, all seven outputs to 1e-5."""
    import numpy as np
    say("C4 1024 x 1026, 10 levels, k = 8, FP32")
    n = 1024
    nx, ny = n, n + 2
    K = 0.5 * P.lognormal_field(nx, ny, 2.0, 32.0, 1)
    om = 1.0 / (3.0 * K + 0.5)
    lam, mu = np.full((nx, ny), 2.0), np.full((nx, ny), 1.0)
    g = P.biot_consolidation("gpu", nx=n, levels=10, steps=2, k_smooth=8, precision="f32/f32")
    g.elasticity.set_lame_fields(lam, mu)
    g.darcy.set_omega_field(om)
    f32 = lambda a: torch.as_tensor(a, dtype=torch.float32)
    o = P.biot_consolidation("oracle", nx=n, levels=10, steps=2, k_smooth=8, precision="f32/f32",
                             lambda_field=f32(lam), mu_field=f32(mu), omega_field=f32(om)[..., None])
    sg = g.initial_state()
    so = o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    outg = g.step(fe, fd, dde, 0)
    outo = o.step(*so, 0)
    errs = P.step_rel(outg, outo, against_state=(4,))
    say("  relative errors", [f"{e:.1e}" for e in errs])
    assert max(errs) < 1e-5, errs


def test_biot_vcycle_4096_matches_oracle():
    """The HBM-resident Biot workload at its size: periodic Biot 4096 x 4096 (manufactured
    forcing, lambda = mu = 1, K = 1), elasticity V-cycle of 6 levels with k = 2 and the
    fused coupling force, FP64; one physical step (one V-cycle and the Darcy step)
    against the oracle, all seven outputs to 1e-12."""
    say("Biot 4096^2, 6 levels, k = 2, one physical step")
    g = P.biot_periodic("gpu", nx=4096, levels=6, steps=1)
    o = P.biot_periodic("oracle", nx=4096, levels=6, steps=1)
    fe, fd, dde = (t.permute(1, 2, 0) for t in g.initial_state())
    outg = g.step(fe, fd, dde, 1)
    outo = o.step(*o.initial_state(), 1)
    errs = P.step_rel(outg, outo)
    say("  relative errors", [f"{e:.1e}" for e in errs])
    assert max(errs) < 1e-12, errs


def test_nonfinite_residual_stops_the_inner_loop():
    # the V-cycle of the reference has no exact coarse solve; at 64^2 with 3 levels the
    # periodic Biot problem diverges (in the oracle too): the loop must stop and say so
    say("biot divergence detection")
    g = P.biot_periodic("gpu", nx=64, levels=3, steps=1)
    g.elasticityTol = 1e-30
    g.elasticityMaxSteps = 5000
    fe, fd, dde = g.initial_state()
    with pytest.raises(FloatingPointError):
        for t in range(50):
            out = g._step_soa(fe, fd, dde, t)
            fe, fd, dde = out[0], out[1], out[4]
