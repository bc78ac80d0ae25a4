"""Benchmark: GLL point-updates/s of the 2D elastic P-SV forward step on one
MI355X.

Workload: 800x600 quads of 50 m (40 km x 30 km), NGLL=5 (480,000 elements,
7,685,601 GLL points), six seeded horizontal layers (vp ~ U[2000, 6000] m/s,
vs = vp/sqrt(3), rho = 310 vp^0.25), 8 Ricker sources (f0 = 5 Hz), 500 surface
receivers sampled every step, dt = 4e-4 s, free top, Stacey elsewhere, FP32.

A step is one full time step (sources, stiffness + Stacey + assembly,
seismograms, divide/corrector/predictor).  Inputs are resident in HBM when the
timed region starts.  value = nglob * steps / wall time.

Prints one JSON line.  Also reports the stiffness kernel against the HBM
roofline (its algorithmic bytes per launch / its mean launch duration measured
with HIP events on the solver's stream) and the CPU oracle timed on a bounded
sample of the same workload on this host.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def tile_kernel_bytes(nspec, layout, nstacey):
    """Algorithmic bytes of one k_tile launch (one time step, FP32):
    * per GLL point of an element a 2-byte tile-local index (50 B/element);
    * geometry: 5 terms per GLL point (500 B/element), or 5 per element
      (20 B) in tiles of affine elements;
    * moduli (lambda+2mu, mu): per GLL point (200 B/element), or per element
      (8 B) in tiles of element-constant material;
    * per point exclusive to a tile: u, v, 1/M read, u, v written: 40 B;
    * per point shared between tiles: u read once (8 B); per partial slot
      (a tile's share of a shared point): 8 B written;
    * Stacey elements: 6 boundary terms and the velocity per point,
      32 B x 25 = 800 B per element."""
    ge = layout.get("geom_elem_elements", 0)
    me = layout.get("mat_elem_elements", 0)
    per_elem = 50 * nspec + 500 * (nspec - ge) + 20 * ge + 200 * (nspec - me) + 8 * me
    return (per_elem + 40 * layout["nexclusive"] + 8 * layout["nshared"]
            + 8 * layout["npartials"] + 800 * nstacey)


def ac_tile_kernel_bytes(nspec, layout, nstacey):
    """Algorithmic bytes of one k_ac_tile launch (acoustic time step, FP32):
    * per element: 2-byte tile-local index per GLL point (50 B), the Stacey
      slot, colour and free-surface mask (9 B);
    * geometry: 5 terms per GLL point (500 B/element), or 5 per element
      (20 B) in tiles of affine elements;
    * 1/rho per GLL point (100 B/element), or per element (4 B) in tiles of
      element-constant material;
    * per point exclusive to a tile: chi, chi_dot, 1/M read, chi, chi_dot
      written: 20 B;
    * per point shared between tiles: chi read once (4 B); per partial slot:
      4 B written;
    * Stacey elements: |n|, edge weight, 1/(rho vp) and the velocity per
      point, 16 B x 25 = 400 B per element."""
    ge = layout.get("geom_elem_elements", 0)
    me = layout.get("mat_elem_elements", 0)
    per_elem = 59 * nspec + 500 * (nspec - ge) + 20 * ge + 100 * (nspec - me) + 4 * me
    return (per_elem + 20 * layout["nexclusive"] + 4 * layout["nshared"]
            + 4 * layout["npartials"] + 400 * nstacey)


def acoustic_leg(nx=800, nz=600, nstep=2000, warmup=100, cpu_steps=20):
    """2D acoustic forward on the layered workload's geometry: the same six
    layers as fluids (vs = 0), acoustic free surface on top, Stacey on the
    other sides, 8 Ricker sources, 500 receivers 100 m below the surface
    (displacement = rho^-1 grad chi and pressure), dt = 4e-4, FP32."""
    import numpy as np
    from specfempp_amd.config import Receivers2D, layered_box
    from specfempp_amd.solver import Context

    cfg = layered_box(nx=nx, nz=nz, nstep=warmup + nstep, acoustic=True)
    cfg.receivers = Receivers2D(x=cfg.receivers.x, z=cfg.receivers.z - 100.0)
    cfg.seismo_types = ("displacement", "pressure")
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, warmup + nstep,
                  seismo_types=cfg.seismo_types)
    ctx.initialize()
    ctx.run_forward(0, warmup)
    ctx.synchronize()
    t = time.perf_counter()
    ctx.run_forward(warmup, nstep)
    ctx.synchronize()
    wall = time.perf_counter() - t
    seis = ctx.seismograms()
    stiff_ms = ctx.time_stiffness(50)
    nstacey = len(set(cfg.mesh.abs_ispec.tolist()))
    tb = ac_tile_kernel_bytes(ctx.nspec, ctx.layout, nstacey)
    out = {"workload": f"2D acoustic forward, {nx}x{nz} quads of 50 m, 6 seeded fluid layers, "
                       f"free top + Stacey on 3 sides, 8 Ricker sources, 500 receivers "
                       f"(displacement, pressure), dt=4e-4, {nstep} steps",
           "nspec": ctx.nspec, "nglob": ctx.nglob, "wall_s": wall,
           "ms_per_step": wall * 1e3 / nstep,
           "point_updates_per_s": ctx.nglob * nstep / wall,
           "roofline": {"bound": "hbm", "kernel": "k_ac_tile<float, AC_STEP>",
                        "achieved": tb / (stiff_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": tb / (stiff_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": tb, "mean_launch_ms": stiff_ms,
                        "layout": dict(ctx.layout)},
           "seismograms_finite": bool(np.isfinite(seis).all())}
    del ctx
    if cpu_steps:
        import oracle
        o = oracle.Oracle2D(cfg.mesh, cfg.sources, Receivers2D(x=cfg.receivers.x[:1],
                                                               z=cfg.receivers.z[:1]),
                            cfg.dt, cpu_steps)
        t = time.perf_counter()
        o.run(cpu_steps)
        sec = time.perf_counter() - t
        out["cpu_baseline"] = {"value": o.nglob * cpu_steps / sec, "unit": "GLL point-updates/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_steps} time steps of the same acoustic workload "
                                         f"(all 8 sources, 1 receiver), FP32 oracle "
                                         f"restatement, set-up excluded, {sec:.1f} s"}
    return out


def coupled_leg(nx=800, nz=600, nstep=2000, warmup=100):
    """Coupled elastic-acoustic forward: the 800x600 layered workload with its
    top 100 element rows turned into water (rho 1020, vp 1500: a 5 km ocean
    over the layered solid), the water's top edges an acoustic free
    surface, Stacey on the other sides, the 8 Ricker sources moved 0.5-1.5
    km below the seabed, 500 receivers 200 m above it (displacement =
    rho^-1 grad chi there), dt = 4e-4."""
    import numpy as np
    from specfempp_amd.config import Receivers2D, layered_box
    from specfempp_amd.mesh import TOP, acoustic_moduli
    from specfempp_amd.solver import Context

    cfg = layered_box(nx=nx, nz=nz, nstep=warmup + nstep)
    m = cfg.mesh
    water_rows = nz // 6
    ids = np.arange(m.nspec)
    wet = ids // nx >= nz - water_rows
    kw, _, rw = acoustic_moduli(np.float32(1020.0), np.float32(1500.0))
    m.kappa[wet], m.mu[wet], m.rho[wet] = kw, 0.0, rw
    top = wet & (ids // nx == nz - 1)
    m.free_ispec = ids[top].astype(np.int32)
    m.free_type = np.full(int(top.sum()), TOP, np.int32)
    seabed = (nz - water_rows) * 50.0
    # the 8 seeded sources moved 0.5-1.5 km below the seabed, receivers
    # 200 m above it, so that waves cross the interface within the run
    cfg.sources.z = seabed - 500.0 - (cfg.sources.z % 1000.0)
    cfg.receivers = Receivers2D(x=cfg.receivers.x, z=np.full(len(cfg.receivers.x),
                                                              seabed + 200.0))
    ctx = Context(m, cfg.sources, cfg.receivers, cfg.dt, warmup + nstep)
    ctx.initialize()
    ctx.run_forward(0, warmup)
    ctx.synchronize()
    t = time.perf_counter()
    ctx.run_forward(warmup, nstep)
    ctx.synchronize()
    wall = time.perf_counter() - t
    seis = ctx.seismograms()
    out = {"workload": f"2D coupled forward, {nx}x{nz} quads of 50 m: top {water_rows} rows "
                       f"water over the 6 seeded layers, free sea surface, Stacey on the "
                       f"other sides, 8 Ricker sources 0.5-1.5 km below the seabed, 500 "
                       f"receivers 200 m above it, dt=4e-4, {nstep} steps",
           "nspec": ctx.nspec, "nglob_both_media": ctx.nglob, "layout": dict(ctx.layout),
           "wall_s": wall, "ms_per_step": wall * 1e3 / nstep,
           "point_updates_per_s": ctx.nglob * nstep / wall,
           "ms_per_step_event": ctx.last_run_ms() / nstep,
           "seismograms_finite": bool(np.isfinite(seis).all()),
           "max_abs_seismogram": float(np.abs(seis).max())}
    del ctx
    return out


def tile3d_kernel_bytes(nspec, layout, nstacey, assemble=False):
    """Algorithmic bytes of one 3D tile pass (k3_tile, one time step, FP32;
    `assemble`: the timing variant that stores a = (f - Ku)/M instead of
    finishing the step, 24 B per exclusive point instead of 60 B):
    * per element: 2-byte tile-local indices (250 B), geometry (10 terms per
      GLL point, 5000 B, or 40 B for an affine element), moduli (2 per GLL
      point, 1000 B, or 8 B), the absorbing-slot index (4 B);
    * the tile's points gathered once: 12 B per exclusive point and per
      partial slot (a tile's share of a shared point);
    * per exclusive point: v, 1/M, u read and u, v written, 3 components
      (60 B); per partial slot: 12 B written;
    * absorbing elements: 7 face terms per point (3500 B) and the velocity."""
    ge = layout.get("geom_elem_elements", 0)
    me = layout.get("mat_elem_elements", 0)
    per_elem = (254 * nspec + 5000 * (nspec - ge) + 40 * ge + 1000 * (nspec - me)
                + 8 * me)
    nx, npart = layout["nexclusive"], layout["npartials"]
    per_ex = 24 if assemble else 60
    return per_elem + 12 * (nx + npart) + per_ex * nx + 12 * npart + 3500 * nstacey


def measured_traffic(path=os.path.join(ROOT, "profiles", "k_tile_layered_800x600.md")):
    """HBM bytes per k_tile launch on the 800x600 workload as recorded from
    the rocprofv3 FETCH_SIZE / WRITE_SIZE passes in the profile summary."""
    try:
        import re
        m = re.search(r"traffic_bytes_per_launch = (\d+)", open(path).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def step_bytes(nspec, nglob):
    """Canonical bytes of one time step: u, v, a read + written and M^-1 read
    once per global point (52 B) plus 800 B per element."""
    return 52 * nglob + 800 * nspec


def cpu_baseline(cfg, nsteps=8):
    """Times the CPU oracle (FP32, 1 thread) on `nsteps` steps of the same
    mesh, sources and medium (one receiver), excluding set-up."""
    import oracle
    from specfempp_amd.config import Receivers2D
    oracle.build()
    rec = Receivers2D(x=cfg.receivers.x[:1], z=cfg.receivers.z[:1])
    o = oracle.Oracle2D(cfg.mesh, cfg.sources, rec, cfg.dt, nsteps)
    t = time.perf_counter()
    o.run(nstep=nsteps)
    dt = time.perf_counter() - t
    return {"value": o.nglob * nsteps / dt, "unit": "GLL point-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{nsteps} time steps of the full 800x600 layered workload "
                      f"(480,000 elements, all 8 sources, 1 receiver), FP32 oracle "
                      f"restatement, set-up excluded, {dt:.1f} s"}


def cpu_baseline_all_cores(cfg, nsteps=120):
    """The same workload on every host core the process may use (capped by
    OMP_NUM_THREADS): the oracle's element work split over OpenMP threads
    within colour groups (no two elements of a group share a point), the
    node updates split likewise.  Reported beside the 1-core figure."""
    import oracle
    from specfempp_amd.config import Receivers2D
    oracle.build()
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    rec = Receivers2D(x=cfg.receivers.x[:1], z=cfg.receivers.z[:1])
    o = oracle.Oracle2D(cfg.mesh, cfg.sources, rec, cfg.dt, nsteps)
    sec, used = o.time_parallel(nsteps, cores)
    return {"value": o.nglob * nsteps / sec, "unit": "GLL point-updates/s", "cores": used,
            "kind": "port",
            "sample": f"{nsteps} time steps of the full 800x600 layered workload, FP32 "
                      f"oracle restatement split over {used} OpenMP threads by element "
                      f"colour, set-up excluded, {sec:.1f} s"}


def adjoint_leg(nfd=20, every=100):
    """Adjoint-gradient workload: the 80x60 homogeneous box (4000x3000 m,
    77,361 points), its Ricker source at (2000 m, 2000 m) and 50 receivers,
    2000 steps of 5e-4 s; observed data from the same box with a 5% Gaussian
    vs anomaly (sigma 300 m) at (2000 m, 1500 m); L2 waveform misfit from the
    homogeneous starting model; checkpoints every 100 steps.  Reports the
    plain forward solve and the gradient (checkpointed forward + adjoint with
    recompute) in HIP-event time, and the FP32 gradient against fourth-order
    central finite differences (relative step 1e-3) of the FP64 misfit in
    `nfd` seeded smooth random directions (all three parameters at once),
    quantized so that every stencil point is an exact FP32 model."""
    import numpy as np
    from specfempp_amd.adjoint import exact_direction, l2_misfit, misfit_and_gradient
    from specfempp_amd.config import adjoint_case, smooth_perturbation
    from specfempp_amd.solver import Context

    cfg, true = adjoint_case()
    m0 = (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)

    def make(prec):
        c = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=prec)
        c.initialize()
        return c

    c32, c64 = make(4), make(8)
    c32.set_model(*true)
    c32.run_forward(0, cfg.nstep)
    c32.synchronize()
    obs = c32.seismograms()[:, 0].copy()
    c32.set_model(*m0)
    # plain forward solve (best of 3) and gradient (best of 3)
    fwd = []
    for _ in range(3):
        c32.run_forward_checkpointed(every)  # resets the fields to rest
        c32.run_forward(0, cfg.nstep)
        c32.synchronize()
        fwd.append(c32.last_run_ms())
    grads = []
    for _ in range(3):
        t = time.perf_counter()
        chi, g32 = misfit_and_gradient(c32, m0, obs, every=every)
        wall = (time.perf_counter() - t) * 1e3
        f_ms, a_ms = c32.last_adjoint_ms()
        grads.append((f_ms + a_ms, f_ms, a_ms, wall))
    best = min(grads)
    fwd_ms = min(fwd)
    # finite-difference checks (FP64 misfit)
    obs64 = obs.astype(np.float64)

    def chi64(model):
        c64.set_model(*model)
        c64.run_forward_checkpointed(every)
        c64.synchronize()
        return l2_misfit(c64.seismograms()[:, 0], obs64, cfg.dt)

    rng = np.random.default_rng(2024)
    base = [x.astype(np.float64) for x in m0]
    errs = []
    for _ in range(nfd):
        dirs = [exact_direction(x32, smooth_perturbation(cfg.mesh, rng) * x * 1e-3)
                for x32, x in zip(m0, base)]

        def at(t):
            return [(x + t * d).astype(np.float32) for x, d in zip(base, dirs)]
        # fourth-order central difference of chi along the direction
        fd = (8.0 * (chi64(at(1)) - chi64(at(-1))) - (chi64(at(2)) - chi64(at(-2)))) / 12.0
        an = sum(float(np.sum(gg * d)) for gg, d in zip(g32, dirs))
        errs.append(abs(an - fd) / abs(fd))
    return {"workload": "80x60 homogeneous box, 77,361 points, 2000 steps of 5e-4 s, "
                        "50 receivers, L2 misfit vs a 5% Gaussian vs anomaly, "
                        f"checkpoints every {every} steps",
            "forward_ms": fwd_ms, "gradient_ms": best[0], "checkpointed_forward_ms": best[1],
            "adjoint_with_recompute_ms": best[2], "gradient_wall_ms": best[3],
            "gradient_over_forward": best[0] / fwd_ms, "misfit": chi,
            "fd_checks": {"directions": nfd, "kind": "smooth random (6 Fourier modes per "
                          "parameter)", "stencil": "fourth-order central",
                          "relative_step": 1e-3, "misfit_precision": "f64",
                          "gradient_precision": "f32", "max_rel_err": max(errs),
                          "median_rel_err": float(np.median(errs))}}


def batch_leg(nshots=64, nstep=3000, baseline_steps=300):
    """Batched multi-shot forward: 64 independent shots, nx ~ U[40, 400] quads
    of 50 m, nz = round(0.75 nx), six seeded layers each, one Ricker source
    (f0 ~ U[5, 15] Hz) and 100 surface receivers per shot, dt = 4e-4 s, 3000
    steps, FP32.  The layout (tile size, graph length, occupancy variant of
    the tile kernel) comes from the autotuner.  All shots advance together, one launch per kernel per step.
    The plain per-shot GPU path (one context per shot, run one after the
    other) is timed on `baseline_steps` steps of every shot and scaled to
    `nstep`."""
    from specfempp_amd.autotune import autotune
    from specfempp_amd.config import batched_shots
    from specfempp_amd.solver import Context

    t = time.perf_counter()
    shots = batched_shots(nshots=nshots, nstep=nstep)
    gen_s = time.perf_counter() - t
    t = time.perf_counter()
    best, timings = autotune(shots, use_cache=False)
    tune_s = time.perf_counter() - t
    t = time.perf_counter()
    ctx = Context.batch(shots, options=best)
    ctx.initialize()
    setup_s = time.perf_counter() - t
    ctx.synchronize()
    t = time.perf_counter()
    ctx.run_forward(0, nstep)
    ctx.synchronize()
    wall = time.perf_counter() - t
    ev_ms = ctx.last_run_ms()
    stiff_ms = ctx.time_stiffness(20)
    nstacey = sum(len(set(c.mesh.abs_ispec.tolist())) for c in shots)
    sb = tile_kernel_bytes(ctx.nspec, ctx.layout, nstacey)
    nglob, nspec, layout = ctx.nglob, ctx.nspec, dict(ctx.layout)
    ctx.close()
    # plain per-shot path: default layout, one context per shot, in turn
    per_shot_ms = 0.0
    for c in shots:
        cc = Context(c.mesh, c.sources, c.receivers, c.dt, baseline_steps)
        cc.initialize()
        cc.run_forward(0, baseline_steps)
        cc.synchronize()
        per_shot_ms += cc.last_run_ms() / baseline_steps
        cc.close()
    return {"workload": f"{nshots} shots, nx ~ U[40, 400], nz = 0.75 nx, 50 m quads, 6 seeded "
                        f"layers, 1 Ricker source and 100 receivers per shot, dt=4e-4, "
                        f"{nstep} steps",
            "nspec": nspec, "nglob": nglob, "layout": layout, "options": best,
            "autotune_ms_per_step": {k: v for k, v in timings.items()},
            "shots_per_s": nshots / wall, "wall_s": wall, "event_ms": ev_ms,
            "ms_per_step": wall * 1e3 / nstep,
            "point_updates_per_s": nglob * nstep / wall,
            "k_tile": {"mean_launch_ms": stiff_ms, "algorithmic_bytes_per_launch": sb,
                       "achieved_GBs": sb / (stiff_ms * 1e-3) / 1e9,
                       "frac": sb / (stiff_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "per_shot_path": {"ms_per_step_sum": per_shot_ms,
                              "shots_per_s": nshots / (per_shot_ms * nstep * 1e-3),
                              "timed_steps_per_shot": baseline_steps},
            "speedup_vs_per_shot": per_shot_ms * nstep * 1e-3 / wall,
            "host_s": {"generate": gen_s, "autotune": tune_s, "assemble": setup_s}}


def hex_leg(n=64, nstep=4000, with_adjoint=True):
    """3D elastic forward: 64^3 hex8 elements of 100 m (6.4 km cube, 262,144
    elements), vp linear 3000 -> 6000 m/s with depth, vs = vp/sqrt(3), rho by
    Gardner, +-2 % seeded Gaussian random perturbation (correlation 500 m);
    free top, Stacey on five faces; explosion moment tensor (1e15), Ricker
    f0 = 2 Hz at 3 km depth; 256 surface receivers; dt = 1.25e-3 s (2e-3 s is
    beyond the stability limit), 4000 steps, FP32.  Then the adjoint
    gradient (kappa, mu, rho) of the L2 misfit against zero data, with
    checkpoints every 100 steps."""
    import numpy as np
    from specfempp_amd.mesh3d import hex_box_case
    from specfempp_amd.solver3d import Context3D

    t = time.perf_counter()
    cfg = hex_box_case(n=n, nstep=nstep)
    gen_s = time.perf_counter() - t
    t = time.perf_counter()
    ctx = Context3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, nstep)
    ctx.initialize()
    setup_s = time.perf_counter() - t
    ctx.synchronize()
    t = time.perf_counter()
    ctx.run_forward(0, nstep)
    ctx.synchronize()
    wall = time.perf_counter() - t
    seis = ctx.seismograms()
    stiff_ms = ctx.time_stiffness(10)
    nglob, nspec, geometry, layout = ctx.nglob, ctx.nspec, ctx.geometry, dict(ctx.layout)
    nstacey = len(set(cfg.mesh.abs_ispec.tolist()))
    tb = tile3d_kernel_bytes(nspec, layout, nstacey, assemble=True)
    adjoint = None
    if with_adjoint:
        # gradient of chi = 1/2 dt sum |s|^2 (the L2 misfit against zero data):
        # checkpointed forward, then recompute + adjoint, one GPU
        every = 100
        t = time.perf_counter()
        ctx.run_forward_checkpointed(every)
        ctx.synchronize()
        s = ctx.seismograms()[:, 0].astype(np.float64)
        ctx.run_adjoint((cfg.dt * s).astype(np.float32))
        ctx.synchronize()
        grad_wall = time.perf_counter() - t
        fwd_ms, adj_ms = ctx.last_adjoint_ms()
        g = ctx.gradient()
        adjoint = {
            "every": every, "checkpointed_forward_ms": fwd_ms, "adjoint_ms": adj_ms,
            "gradient_ms": fwd_ms + adj_ms, "gradient_wall_s": grad_wall,
            "gradient_over_forward": (fwd_ms + adj_ms) / (wall * 1e3),
            "gradient_finite": bool(np.isfinite(g).all()),
            "gradient_norms": [float(np.linalg.norm(g[i])) for i in range(3)],
            "device_buffers_gb": (2 * ((nstep + every - 1) // every) + 2 * every)
                                 * 3 * nglob * 4 / 1e9}
    ctx.close()
    canon = 76 * nglob + 6500 * nspec
    return {"workload": f"3D elastic, {n}^3 hex8 of 100 m, vp 3000->6000 m/s +-2 % random, "
                        f"free top + Stacey on 5 faces, explosion source, 256 receivers, "
                        f"dt={cfg.dt}, {nstep} steps",
            "nspec": nspec, "nglob": nglob, "wall_s": wall, "ms_per_step": wall * 1e3 / nstep,
            "point_updates_per_s": nglob * nstep / wall,
            "canonical_bytes_per_step": canon,
            "canonical_frac": canon * nstep / wall / 1e9 / HBM_PEAK_GBS,
            "tile_pass": {"kernel": "k3_tile<float, assemble> (timing variant: a = "
                                    "(f - K u - B v)/M stored, no Newmark update)",
                          "mean_launch_ms": stiff_ms,
                          "algorithmic_bytes_per_launch": tb,
                          "achieved_GBs": tb / (stiff_ms * 1e-3) / 1e9,
                          "frac": tb / (stiff_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "layout": layout},
            "seismograms_finite": bool(np.isfinite(seis).all()),
            "geometry": geometry, "adjoint": adjoint,
            "host_s": {"generate": gen_s, "assemble": setup_s}}


def _rel(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def errors_leg(c1_oracle_steps=200):
    """Accuracy at the full workload sizes (run with --only-leg errors):
    every workload's FP32 seismograms against an FP64 run of the same GPU
    code, plus the FP32 oracle where it finishes in seconds to minutes:
    config 80x60 (all 2000 steps, seismograms), 800x600 (displacement field
    at every GLL point after `c1_oracle_steps` steps, compared point by point
    through both numberings), the two smallest batched shots (all 3000 steps,
    seismograms).  The 3D oracle's brute-force point location makes the
    64^3 set-up take hours; its parity is tested on the same generator at
    12^3."""
    import numpy as np
    import oracle
    from specfempp_amd.adjoint import misfit_and_gradient
    from specfempp_amd.config import adjoint_case, batched_shots, homogeneous_box, layered_box
    from specfempp_amd.mesh3d import hex_box_case
    from specfempp_amd.solver import Context, simulate, simulate_batch
    from specfempp_amd.solver3d import simulate3d

    out = {}
    # 80x60 homogeneous box, 2000 steps
    cfg = homogeneous_box()
    s32, _ = simulate(cfg)
    s64, _ = simulate(cfg, precision=8)
    o = oracle.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    so = o.run()
    out["80x60"] = {"fp32_vs_fp64": _rel(s32, s64), "fp32_vs_oracle_fp32": _rel(s32[:, 0], so[:, 0]),
                    "quantity": "displacement seismograms, 50 receivers, 2000 steps"}
    # 800x600 layered, 5000 steps
    cfg = layered_box()
    s32, c32 = simulate(cfg)
    s64, _ = simulate(cfg, precision=8)
    from specfempp_amd.config import Receivers2D
    n = c1_oracle_steps
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, n)
    ctx.initialize()
    ctx.run_forward(0, n)
    ctx.synchronize()
    ug = ctx.get_field("u")
    ibg = ctx.ibool().reshape(-1)
    o = oracle.Oracle2D(cfg.mesh, cfg.sources, Receivers2D(x=cfg.receivers.x[:1],
                        z=cfg.receivers.z[:1]), cfg.dt, n)
    o.run(nstep=n)
    uo = o.get("u").reshape(-1, 2)
    ibo = o.get("ibool").astype(np.int64).reshape(-1)
    out["800x600"] = {"fp32_vs_fp64": _rel(s32, s64),
                      "quantity": "displacement seismograms, 500 receivers, 5000 steps",
                      "field_fp32_vs_oracle_fp32": _rel(ug[ibg], uo[ibo]),
                      "field_quantity": f"displacement at all 12,000,000 local GLL points "
                                        f"after {n} steps"}
    del c32, ctx, o
    # adjoint case: gradient and seismograms, FP32 vs FP64
    cfg, true = adjoint_case()
    m0 = (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)
    res = {}
    for prec in (4, 8):
        c = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=prec)
        c.initialize()
        c.set_model(*true)
        c.run_forward(0, cfg.nstep)
        c.synchronize()
        obs = c.seismograms()[:, 0].astype(np.float64)
        res[prec] = misfit_and_gradient(c, m0, obs.astype(c.dtype), every=100)
        res[prec] = (res[prec][1], c.seismograms()[:, 0].copy())
    out["adjoint_80x60"] = {"gradient_fp32_vs_fp64": [_rel(res[4][0][i], res[8][0][i])
                                                      for i in range(3)],
                            "seismograms_fp32_vs_fp64": _rel(res[4][1], res[8][1])}
    # 64 batched shots
    shots = batched_shots()
    b32, _ = simulate_batch(shots)
    b64, _ = simulate_batch(shots, precision=8)
    small = sorted(range(len(shots)), key=lambda k: shots[k].mesh.nspec)[:2]
    orc = []
    for k in small:
        c = shots[k]
        so = oracle.Oracle2D(c.mesh, c.sources, c.receivers, c.dt, c.nstep).run()
        orc.append(_rel(b32[k][:, 0], so[:, 0]))
    # shots whose receivers see an arrival within the 3000 steps (the others
    # record only the round-off tail of the discrete operator)
    energy = np.array([float(np.sum(np.asarray(b, np.float64) ** 2)) for b in b64])
    live = [k for k in range(len(shots)) if energy[k] >= 1e-6 * energy.max()]
    out["64_shots"] = {"fp32_vs_fp64": _rel(np.concatenate([x.ravel() for x in b32]),
                                            np.concatenate([x.ravel() for x in b64])),
                       "shots_with_arrivals": len(live),
                       "worst_shot_with_arrivals_fp32_vs_fp64":
                           max(_rel(b32[k], b64[k]) for k in live),
                       "per_shot": [{"shot": k, "nspec": int(shots[k].mesh.nspec),
                                     "energy_over_max": float(energy[k] / energy.max()),
                                     "fp32_vs_fp64": _rel(b32[k], b64[k])}
                                    for k in range(len(shots))],
                       "smallest_two_shots_fp32_vs_oracle_fp32": orc}
    # 64^3 hex, 4000 steps
    cfg = hex_box_case()
    h32, c = simulate3d(cfg)
    del c
    h64, c = simulate3d(cfg, precision=8)
    del c
    out["hex64"] = {"fp32_vs_fp64": _rel(h32, h64),
                    "quantity": "displacement seismograms, 256 receivers, 4000 steps"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nx", type=int, default=800)
    ap.add_argument("--nz", type=int, default=600)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint-gradient leg (80x60 box, 2000 steps)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-shots leg (64 shots, 3000 steps)")
    ap.add_argument("--no-3d", action="store_true",
                    help="skip the 3D leg (64^3 hex, 4000 steps)")
    ap.add_argument("--no-acoustic", action="store_true",
                    help="skip the acoustic leg (800x600 fluid layers, 2000 steps)")
    ap.add_argument("--no-coupled", action="store_true",
                    help="skip the coupled elastic-acoustic leg (800x600, 2000 steps)")
    ap.add_argument("--only-leg", choices=("adjoint", "batch", "hex3d", "errors", "acoustic",
                                           "coupled"),
                    default=None,
                    help="profiling aid: run only this leg and print its JSON")
    ap.add_argument("--hex-steps", type=int, default=4000,
                    help="time steps of the 3D leg (4000 is the workload; fewer only "
                         "to profile it)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per k_tile launch measured with rocprofv3 --pmc "
                         "(FETCH_SIZE x 2 + WRITE_SIZE, copy-kernel calibrated), if known")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB (the "
                         "known-size kernel that calibrates rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_probe_mb > 0:
        import torch
        n = args.copy_probe_mb * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                          "launches": 5}))
        return
    if args.only_leg:
        leg = {"adjoint": adjoint_leg, "batch": batch_leg,
               "hex3d": lambda: hex_leg(nstep=args.hex_steps),
               "errors": errors_leg,
               "acoustic": lambda: acoustic_leg(cpu_steps=0 if args.no_cpu_baseline
                                                else 20),
               "coupled": coupled_leg}[args.only_leg]
        print(json.dumps({args.only_leg: leg()}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: sharding the mesh over "
                                   "several GPUs (tiles per rank, halo exchange of the "
                                   "shared points) is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from specfempp_amd.config import layered_box
    from specfempp_amd.solver import Context

    nstep = args.warmup + args.steps
    cfg = layered_box(nx=args.nx, nz=args.nz, nstep=nstep)
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, nstep,
                  seismo_types=cfg.seismo_types)
    ctx.initialize()
    nstacey = len(set(cfg.mesh.abs_ispec.tolist()))

    if args.warmup > 0:
        ctx.run_forward(0, args.warmup)
        ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.run_forward(args.warmup, args.steps)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ctx.last_run_ms()

    seis = ctx.seismograms()
    import numpy as np
    finite = bool(np.isfinite(seis).all())
    import ctypes
    gm, mis = ctypes.c_int(), ctypes.c_longlong()
    ctx.lib.smx_ctx_geometry(ctx.h, ctypes.byref(gm), ctypes.byref(mis))
    geometry = {"mode": ("point", "element", "nodes")[gm.value], "mismatches": mis.value}

    stiff_ms = ctx.time_stiffness(50)
    sb = tile_kernel_bytes(ctx.nspec, ctx.layout, nstacey)
    achieved = sb / (stiff_ms * 1e-3) / 1e9

    value = ctx.nglob * args.steps / wall
    out = {
        "metric": "GLL point-updates/s (FP32, 1 GPU)",
        "value": value,
        "unit": "GLL point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"2D elastic P-SV forward, {args.nx}x{args.nz} quads of 50 m, "
                               f"NGLL=5, 6 seeded layers, 8 Ricker sources, 500 receivers, "
                               f"dt=4e-4",
                   "nspec": ctx.nspec, "nglob": ctx.nglob, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_tile<float, MODE_STEP>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": args.traffic if args.traffic is not None else (
                         measured_traffic() if (args.nx, args.nz) == (800, 600) else None),
                     "algorithmic_bytes_per_launch": sb, "mean_launch_ms": stiff_ms,
                     "layout": ctx.layout},
        "step_roofline": {"canonical_bytes_per_step": step_bytes(ctx.nspec, ctx.nglob),
                          "achieved_GBs": step_bytes(ctx.nspec, ctx.nglob) * args.steps
                          / wall / 1e9,
                          "frac": step_bytes(ctx.nspec, ctx.nglob) * args.steps / wall / 1e9
                          / HBM_PEAK_GBS,
                          "event_ms": ev_ms},
        "seismograms_finite": finite,
        "geometry": geometry,
    }
    del ctx
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_batch:
        out["batch"] = batch_leg()
    if not args.no_3d:
        out["hex3d"] = hex_leg(nstep=args.hex_steps)
    if not args.no_acoustic:
        out["acoustic"] = acoustic_leg(cpu_steps=0 if args.no_cpu_baseline else 20)
    if not args.no_coupled:
        out["coupled"] = coupled_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
