"""Layout autotuner: picks the tile size (elements per workgroup), the HIP
graph length and the occupancy variant of the tile kernel for a workload by
timing the fused forward step.

The reference has no counterpart: Kokkos picks team sizes with AUTO and the
chunk size is a compile-time constant (ParallelConfig in
compute_stiffness_interaction.tpp).  Here the choice matters in two regimes:

* small meshes (the 80x60 box has 75 tiles of 64 elements for 256 CUs) are
  latency-bound: smaller tiles spread the step over more CUs and longer
  graphs amortise launch overhead;
* large meshes and batches are HBM-bound: 64-element tiles (8x8 squares)
  minimise the points shared between tiles and win.

Candidates are the tile sizes {64, 32, 16} x the graph lengths {8, 32} x
the occupancy variants of the FP32 time-step tile kernel (7, 8 and 6 waves
per SIMD; FP64 contexts have one), at every mesh size.  Each tile size is
assembled once; its graph lengths and occupancy variants are then switched
on that context (smx_ctx_set_launch re-captures and uploads the step graph
before the timed window).  Every candidate is timed with HIP events on the
context's own stream over the same steps: a warm-up run, then one
predictor, a body whose length is a multiple of every graph length (so each
candidate replays only its graph), and the final step.  Results are cached
per workload signature.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

from .solver import Context

TILE_CHOICES = (64, 32, 16)
GRAPH_CHOICES = (8, 32)
# occupancy variants of k_tile<float, MODE_STEP>: 7 (the default), 8, 6
WPE_CHOICES = (7, 8, 6)
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
    h.update(f"p{precision}v3".encode())
    return h.hexdigest()[:16]


def candidates(configs, precision=4):
    """Every tile size x every graph length x every occupancy variant, at
    every mesh size (FP64 contexts have the default variant only)."""
    wpes = WPE_CHOICES if precision == 4 else (0,)
    return [{"tile_elems": t, "graph_steps": g, "waves_per_eu": w}
            for t in TILE_CHOICES for g in GRAPH_CHOICES for w in wpes]


def label(opt):
    """tile{t}_graph{g}_wpe{w}; "default" for the kernel's own occupancy
    (waves_per_eu 0: 7 waves per SIMD in FP32, unconstrained in FP64)."""
    w = opt.get("waves_per_eu", 0)
    return f"tile{opt['tile_elems']}_graph{opt['graph_steps']}_" + (f"wpe{w}" if w else "default")


def _context(configs, tile, precision, nstep):
    options = {"tile_elems": tile}
    if len(configs) == 1:
        c = configs[0]
        return Context(c.mesh, c.sources, c.receivers, c.dt, nstep, nsbs=c.nsbs,
                       seismo_types=c.seismo_types, precision=precision, options=options)
    shots = [_with_nstep(c, nstep) for c in configs]
    return Context.batch(shots, precision=precision, options=options)


def _time_tile(configs, tile, opts, precision, steps, warmup, verbose):
    """Times every candidate of one tile size on one assembly; returns
    {label: ms per step}."""
    per = warmup + steps
    ctx = _context(configs, tile, precision, per * len(opts))
    out = {}
    try:
        ctx.initialize()
        for k, opt in enumerate(opts):
            t = time.perf_counter()
            ctx.set_launch(opt["graph_steps"], opt["waves_per_eu"])
            s0 = k * per
            ctx.run_forward(s0, warmup)
            ctx.synchronize()
            ctx.run_forward(s0 + warmup, steps)
            ctx.synchronize()
            out[label(opt)] = ctx.last_run_ms() / steps
            if verbose:
                print(f"autotune {label(opt)}: {out[label(opt)] * 1e3:.2f} us/step "
                      f"({time.perf_counter() - t:.1f} s)", flush=True)
    finally:
        ctx.close()
    return out


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
    cands = candidates(configs, precision)
    for tile in TILE_CHOICES:
        opts = [c for c in cands if c["tile_elems"] == tile]
        timings.update(_time_tile(configs, tile, opts, precision, steps, warmup, verbose))
    best = min(cands, key=lambda c: timings[label(c)])
    if use_cache:
        cache[key] = {"best": best, "timings": timings}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best, timings
