"""The parameter sweeps of the reference's run scripts, run in one process.

    python xlb-biot_amd/examples/sweeps.py periodic            # runPeriodic.sh
    python xlb-biot_amd/examples/sweeps.py periodic_mg         # runPeriodicMG.sh
    python xlb-biot_amd/examples/sweeps.py periodic_mg_altgrad # runPeriodicMG_altgrad.sh
    python xlb-biot_amd/examples/sweeps.py terzaghi            # runTerzaghi.sh

``--coarse`` adds the scripts' first pass (fewer refinements, every coupling; the
scripts run it when given any argument), ``--refine R`` caps the refinements,
``--dry-run`` lists the studies, ``--limit N`` runs the first N.  The scripts
start the coarse pass's studies as concurrent processes; on one GPU they run one
after another.  Every study
writes (or reuses) the drivers' ``.npz`` files, so an interrupted sweep resumes
where it stopped.
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import periodic  # noqa: E402
import terzaghi  # noqa: E402

COUPLINGS = ("explicit", "implicit", "centered")


def _seq(a, b, step):
    """``seq a step b`` of the scripts, as rounded floats."""
    n = int(round((b - a) / step)) + 1
    return [round(a + i * step, 10) for i in range(n)]


def plan(name, coarse=False):
    """``(driver, kwargs)`` of every study of a sweep, in the script's order."""
    runs = []
    if name == "periodic":
        # runPeriodic.sh: n0 = 10, alpha = 0.5, relative elasticity steps 0.1 .. 1.5
        for refine, passes in ((4, coarse), (9, True)):
            if not passes:
                continue
            for v in _seq(0.1, 1.5, 0.1):
                for c in COUPLINGS:
                    runs.append((periodic, dict(nx=10, refinements=refine, alpha=0.5, coupling=c, elastic_steps=v)))
    elif name in ("periodic_mg", "periodic_mg_altgrad"):
        # runPeriodicMG(_altgrad).sh: n0 = 4, automatic depth, 1 .. 6 V-cycles per step
        g = name.endswith("altgrad")
        for refine, passes in ((3, coarse), (5, True)):
            if not passes:
                continue
            for alpha in _seq(0.3, 1.0, 0.1):
                for v in range(1, 7):
                    for c in COUPLINGS:
                        if refine == 5 and c == "explicit" and alpha >= (0.9 if g else 0.8):
                            continue
                        if refine == 5 and c == "implicit" and not g and alpha >= 0.9:
                            continue
                        runs.append((periodic, dict(nx=4, levels=-1, refinements=refine, alpha=alpha, coupling=c,
                                                    elastic_steps=v, gradient=g)))
    elif name == "terzaghi":
        # runTerzaghi.sh: n0 = 4, automatic depth, alpha 1.0 / 0.9 / 0.8, 1 .. 6 V-cycles
        for refine, passes in ((3, coarse), (5, True)):
            if not passes:
                continue
            for alpha in (1.0, 0.9, 0.8):
                for v in range(1, 7):
                    for c in COUPLINGS:
                        runs.append((terzaghi, dict(nx=4, levels=-1, refinements=refine, alpha=alpha, coupling=c,
                                                    elastic_steps=v)))
    else:
        raise ValueError(f"unknown sweep {name!r}")
    return runs


def run(name, coarse=False, refine=None, dry_run=False, output_root=".", limit=None):
    """Run a sweep (its first ``limit`` studies); returns ``[(driver name, kwargs,
    {nx: record})]``."""
    out = []
    for mod, kw in plan(name, coarse)[:limit]:
        kw = dict(kw)
        if refine is not None:
            kw["refinements"] = min(kw["refinements"], refine)
        kw["output_dir"] = os.path.join(output_root, os.path.relpath(mod.OUTPUT_DIR, "."))
        label = f"{mod.__name__} " + " ".join(f"{k}={v}" for k, v in kw.items() if k != "output_dir")
        print(label, flush=True)
        if dry_run:
            continue
        res = mod.study(**kw, plots=False, print_info_rate=0)
        out.append((mod.__name__, kw, res))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Parameter sweeps of the poroelastic studies (run*.sh).")
    ap.add_argument("sweep", choices=["periodic", "periodic_mg", "periodic_mg_altgrad", "terzaghi"])
    ap.add_argument("--coarse", action="store_true", help="also run the scripts' first, coarse pass")
    ap.add_argument("--refine", type=int, default=None, help="cap the refinements of every study")
    ap.add_argument("--dry-run", action="store_true", help="list the studies only")
    ap.add_argument("--limit", type=int, default=None, help="run the first LIMIT studies only")
    ap.add_argument("--output-root", default=".", help="folder holding results/poroelasticity/...")
    a = ap.parse_args(argv)
    res = run(a.sweep, a.coarse, a.refine, a.dry_run, a.output_root, a.limit)
    if res:
        rows = [(k["alpha"], k["coupling"], k["elastic_steps"], n, r) for _, k, rr in res for n, r in rr.items()]
        key = "rel_err_sub" if a.sweep == "terzaghi" else "rel_err_l2_eta"
        print(f"{'alpha':>6} {'coupling':>9} {'e':>5} {'nx':>5} {key:>16}")
        for alpha, c, e, n, r in rows:
            print(f"{alpha:>6} {c:>9} {e:>5} {n:>5} {float(np.asarray(r[key])):>16.6e}")


if __name__ == "__main__":
    main()
