"""Launch-shape autotuner of the sweep kernels.

The sweep kernel's launch shape is ``rows``: how many grid rows one wavefront
walks (a power of two in [1, 64]; at least 2 for the residual sweep, whose
restriction pairs rows).  Few rows give more wavefronts (parallelism,
shorter tails) but re-read more halo rows from L2; many rows amortise the halo
and the pipeline prologue.  The best value depends on grid size, lattice and
dtype, so it is measured: each candidate runs a few graph-captured launches timed
with HIP events, and the fastest is cached per key in a JSON file
(``$XB_TUNE_CACHE``, default ``~/.cache/xlb_biot/tune.json``).

The multigrid hierarchy depth is tuned the same way: the reference V-cycle (sum
restriction, injection, smoothing only on the coarsest grid) contracts for some
Lame parameters only with few levels and diverges with more (for mu >= 1 on the
periodic D2Q8 problem, with three levels or more), so :func:`tune_levels`
runs a few V-cycles of the actual problem at each candidate depth and keeps the
depth with the shortest predicted time to the tolerance.
"""
import json
import math
import os
import time

import torch

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)
_cache = None


def cache_path():
    return os.environ.get("XB_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "xlb_biot", "tune.json"))


def _load():
    global _cache
    if _cache is None:
        try:
            with open(cache_path()) as fh:
                _cache = json.load(fh)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def lookup(key):
    """Tuned ``rows`` for ``key`` (0 = built-in choice when never tuned)."""
    return int(_load().get(key, {}).get("rows", 0))


def store(key, rows, timings):
    c = _load()
    c[key] = {"rows": int(rows), "us": {str(k): round(v * 1e6, 3) for k, v in timings.items()}}
    try:
        os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
        with open(cache_path(), "w") as fh:
            json.dump(c, fh, indent=1, sort_keys=True)
    except OSError:
        pass


def time_launch(launch, reps=20):
    """Average seconds per ``launch()`` over ``reps`` back-to-back launches."""
    for _ in range(3):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def tune(key, launch_with_rows, candidates=CANDIDATES, reps=20):
    """Time ``launch_with_rows(rows)`` for every candidate, cache and return the best."""
    timings = {r: time_launch(lambda: launch_with_rows(r), reps) for r in candidates}
    best = min(timings, key=timings.get)
    store(key, best, timings)
    return best, timings


# Plain-step launch kinds (the value cached under a solver's "/multi" key): one step
# per sweep, two per sweep (register temporal blocking), or m steps on LDS tiles.
SWEEP1, SWEEP2 = 1, 2
# tile launches: TILE_BASE * k + m steps per launch, k = 1 the library's tile shape,
# k = 2: 32 x 32 tiles in blocks of 256 threads, k = 3: 32 x 32 in blocks of 512,
# k = 4: 44 x 48 tiles in blocks of 704 threads
TILE_BASE = 100
TILE_STEPS = (2, 3, 4, 6, 8)
TILE_THREADS = (256, 512, 704)
TILE_SHAPES = {256: (32, 32), 512: (32, 32), 704: (44, 48)}


def tile_choice(m, threads=0):
    k = {0: 1, 256: 2, 512: 3, 704: 4}[int(threads)]
    return TILE_BASE * k + int(m)


def tile_blocks(nx, ny, m, threads):
    """Blocks of a tile launch of ``m`` steps with ``threads`` threads per block."""
    tx, ty = TILE_SHAPES[threads]
    return -(-nx // (tx - 2 * m)) * -(-ny // (ty - 2 * m))


def library_tile_threads(nx, ny, m):
    """The library's own tile shape for ``threads`` 0: 512 threads for grids of at
    most four 32 x 32 tiles per CU (256 CUs), else 256."""
    return 512 if tile_blocks(nx, ny, m, 512) <= 4 * 256 else 256


def steps_per_launch(choice):
    """Steps one launch of ``choice`` advances."""
    return choice % TILE_BASE if choice > TILE_BASE else choice


def tile_threads(choice):
    """Threads per tile block of a tile ``choice`` (0: the library's choice)."""
    return (0, 0, 256, 512, 704)[choice // TILE_BASE]


def tune_multi(key, launches, reps=20):
    """``launches``: {choice: launch()}; times each, keeps the choice with the least
    time per step and caches it under ``key`` (see :func:`lookup`)."""
    timings = {c: time_launch(fn, reps) / steps_per_launch(c) for c, fn in launches.items()}
    best = min(timings, key=timings.get)
    store(key, best, timings)
    return best, timings


def cycles_to_tol(norms, tol, tail=20):
    """Cycles a residual history needs to reach ``tol`` times its first value: the
    measured count when it gets there, else an extrapolation at the geometric rate
    of its last ``tail`` cycles (inf when that rate is not below one)."""
    r = [float(x) for x in norms]
    r0 = r[0]
    if not (r0 == r0) or r0 == float("inf"):
        return float("inf")
    if r0 == 0.0:
        return 1.0
    for k, v in enumerate(r):
        if not (v == v) or v == float("inf"):
            return float("inf")
        if v <= tol * r0:
            return float(k + 1)
    tail = min(tail, len(r) - 1)
    a, b = r[-1 - tail], r[-1]
    if a <= 0.0 or b >= a:
        return float("inf")
    rate = (b / a) ** (1.0 / tail)
    return len(r) + math.log(tol * r0 / b) / math.log(rate)


def probe_vcycle(vcycle_graph, state, cycles=60):
    """Residual norms of ``cycles`` replays of a V-cycle graph from ``state``
    (recorded on the device, read once) and the seconds per cycle."""
    vcycle_graph.load(state)
    hist = torch.empty(cycles, dtype=torch.float64, device=state.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(cycles):
        vcycle_graph.replay()
        hist[k].copy_(vcycle_graph.norm)
    out = hist.cpu().tolist()
    return out, (time.perf_counter() - t0) / cycles


def tune_levels(key, probe, candidates, tol):
    """Depth in ``candidates`` with the shortest predicted time to reach ``tol``
    (cycles to tolerance, measured or extrapolated, times the cost of a cycle);
    ``probe(L)`` returns ``(residual history, seconds per cycle)`` at depth L.
    Deeper candidates stop being tried after one that cannot reach the tolerance
    following one that can.  Cached under ``key``."""
    c = _load()
    if key in c and "levels" in c[key]:
        return int(c[key]["levels"]), c[key].get("predicted_s", {})
    scores = {}
    best = None
    for L in candidates:
        hist, sec = probe(L)
        scores[L] = cycles_to_tol(hist, tol) * sec
        if best is None or scores[L] < scores[best]:
            best = L
        elif scores[L] == float("inf") and scores[best] < float("inf"):
            break
    c[key] = {"levels": int(best), "predicted_s": {str(k): (round(v, 6) if v != float("inf") else "inf")
                                                   for k, v in scores.items()}}
    try:
        os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
        with open(cache_path(), "w") as fh:
            json.dump(c, fh, indent=1, sort_keys=True)
    except OSError:
        pass
    return best, scores
