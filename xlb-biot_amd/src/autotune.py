"""Launch-shape autotuner of the sweep kernels.

The sweep kernel's launch shape is ``rows``: how many grid rows one wavefront
walks (a power of two in [2, 64]).  Few rows give more wavefronts (parallelism,
shorter tails) but re-read more halo rows from L2; many rows amortise the halo
and the pipeline prologue.  The best value depends on grid size, lattice and
dtype, so it is measured: each candidate runs a few graph-captured launches timed
with HIP events, and the fastest is cached per key in a JSON file
(``$XB_TUNE_CACHE``, default ``~/.cache/xlb_biot/tune.json``).
"""
import json
import os

import torch

CANDIDATES = (2, 4, 8, 16, 32, 64)
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
