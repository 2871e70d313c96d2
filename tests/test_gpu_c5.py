"""The batched-Biot workload at full size: 64 periodic Biot instances with square
grids of 64, 128, 256 and 512 cells per side (sizes, Lame parameters and
log-normal permeabilities drawn with seed 3), FP64, eight smoothing steps per
level, inner V-cycles to 1e-8 of the zero state's residual.

* The batched solve against the same 64 instances solved one after another:
  states bit for bit and cycle counts exactly (both run the same kernels with
  the same operands; the device-side convergence test sums the same partials in
  the same order as the host-side one).
* Two instances of the set against the oracle: a fixed number of V-cycles per
  physical step, every output of the coupled step to 1e-12."""
import pytest
import torch

import problems as P

pytestmark = pytest.mark.gpu

# hierarchy depth per grid size (the depths the depth tuner picks for most of the set)
LEVELS = {64: 2, 128: 2, 256: 3, 512: 4}


def say(*a):
    print(*a, flush=True)


def _instance(kind, spec, K, tol=1e-8):
    return P.biot_c5(kind, spec, LEVELS[spec["n"]], tol, 3000, K=K, k_smooth=8)


def test_c5_batch_matches_sequential_bitwise():
    from src.batch import BiotBatch
    specs = P.c5_specs(64, 3)
    fields = [P.c5_permeability(s) for s in specs]
    say("C5: 64 instances, sizes", sorted({s["n"] for s in specs}))
    seq = [_instance("gpu", s, K) for s, K in zip(specs, fields)]
    states = [p.initial_state() for p in seq]
    bb = BiotBatch([_instance("gpu", s, K) for s, K in zip(specs, fields)])
    bstates = bb.initial_state()
    for t in range(2):
        ref = [p._step_soa(fe, fd, dde, t) for p, (fe, fd, dde) in zip(seq, states)]
        got = bb.step(bstates, t)
        say("physical step", t, "cycles", [p.inner_counts[-1] for p in seq])
        for i in range(len(specs)):
            assert bb.problems[i].inner_counts[-1] == seq[i].inner_counts[-1], (t, i)
            assert seq[i].inner_residuals[-1] <= 1e-8, (t, i, seq[i].inner_residuals[-1])
            for k, (a, b) in enumerate(zip(got[i], ref[i])):
                assert torch.equal(a, b), (t, i, k, (a - b).abs().max().item())
        states = [(o[0], o[1], o[4]) for o in ref]
        bstates = [(o[0], o[1], o[4]) for o in got]


@pytest.mark.parametrize("which", [0, 1])
def test_c5_instances_match_oracle(which):
    specs = P.c5_specs(64, 3)
    # the first two 64^2 instances of the set (the oracle runs them on the CPU)
    idx = [i for i, s in enumerate(specs) if s["n"] == 64][which]
    spec = specs[idx]
    K = P.c5_permeability(spec)
    say("C5 instance", idx, spec, "against the oracle")
    g = _instance("gpu", spec, K, tol=None)
    o = _instance("oracle", spec, K, tol=None)
    for p in (g, o):
        p.elasticityTol = None
        p.elasticitySteps = 6
    sg, so = g.initial_state(), o.initial_state()
    fe, fd, dde = (t.permute(1, 2, 0) for t in sg)
    oe, od, odde = so
    for t in range(2):
        outg = g.step(fe, fd, dde, t)
        outo = o.step(oe, od, odde, t)
        errs = P.step_rel(outg, outo)
        say("  step", t, "relative errors", [f"{e:.1e}" for e in errs])
        assert max(errs) < 1e-12, (t, errs)
        fe, fd, dde = outg[0], outg[1], outg[4]
        oe, od, odde = outo[0], outo[1], outo[4]
