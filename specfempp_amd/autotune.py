"""Layout autotuner: picks the tile size (elements per workgroup) and the HIP
graph length for a workload by timing the fused forward step.

The reference has no counterpart: Kokkos picks team sizes with AUTO and the
chunk size is a compile-time constant (ParallelConfig in
compute_stiffness_interaction.tpp).  Here the choice matters in two regimes:

* small meshes (the 80x60 box has 75 tiles of 64 elements for 256 CUs) are
  latency-bound: smaller tiles spread the step over more CUs and longer
  graphs amortise launch overhead;
* large meshes and batches are HBM-bound: 64-element tiles (8x8 squares)
  minimise the points shared between tiles and win.

Candidates are timed with HIP events on the context's own stream.  Every
context captures and instantiates its step graph when it is initialised, and
every candidate is timed over the same steps: one predictor, a body whose
length is a multiple of every graph length (so each candidate replays only
its graph), and the final step.  Results are cached per workload signature
and device.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

from .solver import Context

TILE_CHOICES = (64, 32, 16)
GRAPH_CHOICES = (8, 32)
# timed body steps: a common multiple of every graph length
BODY_STEPS = 64


def _cache_path():
    return os.environ.get("SMX_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache",
                                       "specfempp_amd_autotune.json"))


def signature(configs, precision):
    """Workload key: element counts, receivers, sources and precision."""
    h = hashlib.sha1()
    for c in configs:
        h.update(f"{c.mesh.nspec}:{c.mesh.ngnod}:{len(c.receivers)}:{len(c.sources)}:"
                 f"{len(c.mesh.abs_ispec)};".encode())
    h.update(f"p{precision}".encode())
    return h.hexdigest()[:16]


def candidates(configs):
    """Every tile size x every graph length, at every mesh size."""
    return [{"tile_elems": t, "graph_steps": g} for t in TILE_CHOICES for g in GRAPH_CHOICES]


def _time(configs, options, precision, steps, warmup):
    n = warmup + steps
    if len(configs) == 1:
        c = configs[0]
        ctx = Context(c.mesh, c.sources, c.receivers, c.dt, n, nsbs=c.nsbs,
                      seismo_types=c.seismo_types, precision=precision, options=options)
    else:
        shots = [_with_nstep(c, n) for c in configs]
        ctx = Context.batch(shots, precision=precision, options=options)
    ctx.initialize()
    ctx.run_forward(0, warmup)
    ctx.synchronize()
    ctx.run_forward(warmup, steps)
    ctx.synchronize()
    ms = ctx.last_run_ms() / steps
    ctx.close()
    return ms


def _with_nstep(cfg, n):
    import copy
    c = copy.copy(cfg)
    c.nstep = n
    return c


def autotune(configs, *, precision=4, steps=BODY_STEPS + 1, warmup=BODY_STEPS // 2 + 1,
             use_cache=True, verbose=False):
    """Returns (best options dict, {label: ms per step}) for a Config2D or a
    list of Config2D run as one batch."""
    if not isinstance(configs, (list, tuple)):
        configs = [configs]
    key = signature(configs, precision)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        return cache[key]["best"], cache[key]["timings"]
    timings = {}
    best, best_ms = None, float("inf")
    for opt in candidates(configs):
        t = time.perf_counter()
        ms = _time(configs, opt, precision, steps, warmup)
        label = f"tile{opt['tile_elems']}_graph{opt['graph_steps']}"
        timings[label] = ms
        if verbose:
            print(f"autotune {label}: {ms * 1e3:.2f} us/step ({time.perf_counter() - t:.1f} s)",
                  flush=True)
        if ms < best_ms:
            best, best_ms = opt, ms
    if use_cache:
        cache[key] = {"best": best, "timings": timings}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best, timings
