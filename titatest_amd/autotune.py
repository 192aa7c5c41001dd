"""Autotuner for the forward's launch choices.

The tunable choices (crt_set_tuning) are the factor kernel's workgroup size
(256 threads -- two workgroups per CU where the band window's LDS allows --
or 512 threads, one workgroup with seven update waves) and the back
substitution's prefetch depth (8 or 16 rows of L in flight per wavefront).
Every choice computes the same algorithm in the same operation order.
``tune`` times each candidate on the model itself (complete forwards of the
handle's conductivities, HIP events around the factor and substitution
stages), keeps the fastest and records it in a JSON cache keyed by the
model's shape (nodes, band width, electrodes, wavenumbers, mixed boundaries),
so later handles of the same shape take the choice without timing.  The
panel width (8 columns) is fixed at compile time: it sets the LDS layout and
the panel wave's registers.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

FACTOR_THREADS = (256, 512)
BACKSOLVE_DEPTHS = (8, 16)


def default_cache() -> Path:
    return Path(os.environ.get("CRT_TUNE_CACHE", Path.home() / ".cache" / "titatest_amd" / "tune.json"))


def shape_key(g) -> str:
    m = g.mesh
    mixed = int(any(int(t) == 11 for t in m.typ))
    return f"n{m.sanz}_mb{m.mb}_E{m.eanz}_K{g.K}_mixed{mixed}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _time_forward(g, sigma, reps: int) -> float:
    """Milliseconds per forward (factor + substitution stages, the minimum
    over ``reps`` forwards)."""
    best = float("inf")
    g.forward(sigma)  # buffers, LDS attributes
    for _ in range(reps):
        g.set_timing(True)
        g.forward(sigma)
        t = g.stage_times()
        g.set_timing(False)
        ms = sum(v[0] for k, v in t.items() if k in ("factor_solve", "backsolve", "forward_sweep", "column_copy"))
        best = min(best, ms)
    return best


def tune(g, sigma, cache: Path | None = None, reps: int = 2, force: bool = False):
    """Choose (factor_threads, backsolve_depth) for handle ``g`` (after
    precal) and apply it.  Returns (choice, timings) with timings[(threads,
    depth)] = ms per forward, empty when the choice came from the cache."""
    cache = default_cache() if cache is None else Path(cache)
    key = shape_key(g)
    known = _load(cache)
    if key in known and not force:
        t, d = known[key]["factor_threads"], known[key]["backsolve_depth"]
        g.set_tuning(t, d)
        return (t, d), {}
    timings = {}
    for t in FACTOR_THREADS:
        for d in BACKSOLVE_DEPTHS:
            g.set_tuning(t, d)
            timings[(t, d)] = _time_forward(g, np.asarray(sigma), reps)
    choice = min(timings, key=timings.get)
    g.set_tuning(*choice)
    known[key] = {"factor_threads": choice[0], "backsolve_depth": choice[1],
                  "ms": {f"{t}x{d}": v for (t, d), v in timings.items()}}
    cache.parent.mkdir(parents=True, exist_ok=True)
    tmp = cache.with_suffix(".tmp")
    tmp.write_text(json.dumps(known, indent=1, sort_keys=True))
    os.replace(tmp, cache)
    return choice, timings
