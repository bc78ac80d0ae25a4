"""Batched shots and layout options on the GPU.

* A batched context advances several independent shots with one launch per
  kernel per step; every shot keeps its own tiling and summation order, so
  its seismograms (and adjoint gradients) equal those of its own context bit
  for bit.
* Tile sizes 16 / 32 / 64 only change summation order: seismograms stay
  within 1e-5 of the oracle.  The graph length and the occupancy variant of
  the tile kernel change nothing, also when switched between runs.
"""

import sys

import numpy as np
import pytest

from specfempp_amd._lib import SmxError
from specfempp_amd.autotune import autotune, candidates, label
from specfempp_amd.config import batched_shots
from specfempp_amd.solver import Context, simulate, simulate_batch
from tests._cases import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _shots(n=5, nstep=200, seed=7):
    return batched_shots(nshots=n, seed=seed, nstep=nstep, nrec=12, nmin=10, nmax=36)


def test_batch_equals_individual_shots_bitwise():
    shots = _shots()
    _say("batch of 5 shots (" + ", ".join(c.name for c in shots) + "), 200 steps")
    got, ctx = simulate_batch(shots)
    assert ctx.nshots == len(shots)
    assert ctx.layout["ntiles"] == sum(
        -(-c.mesh.nspec // 64) for c in shots)  # every shot ends its own tiles
    for cfg, s in zip(shots, got):
        ref, _ = simulate(cfg)
        assert s.shape == ref.shape
        assert np.array_equal(s, ref), cfg.name


def test_batch_shot_matches_oracle(oracle_lib):
    shots = _shots(n=3, nstep=300, seed=11)
    _say("batch of 3 shots vs oracle, 300 steps")
    got, _ = simulate_batch(shots)
    for cfg, s in zip(shots, got):
        o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
        ref = o.run(nsbs=cfg.nsbs, types=cfg.seismo_types)
        err = rel_l2(s, ref)
        _say(f"  {cfg.name}: rel L2 vs oracle {err:.3e}")
        assert err <= 1e-5


@pytest.mark.parametrize("tile_elems", [16, 32])
def test_tile_size_option(oracle_lib, tile_elems):
    cfg = _shots(n=1, nstep=300, seed=3)[0]
    _say(f"tile_elems={tile_elems}: {cfg.name}, 300 steps")
    s64, c64 = simulate(cfg)
    s, c = simulate(cfg, options={"tile_elems": tile_elems})
    assert c.layout["ntiles"] == -(-cfg.mesh.nspec // tile_elems)
    assert c.layout["ntiles"] > c64.layout["ntiles"]
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ref = o.run(nsbs=cfg.nsbs, types=cfg.seismo_types)
    assert rel_l2(s, ref) <= 1e-5
    assert rel_l2(s, s64) <= 1e-5


def test_graph_length_changes_nothing():
    cfg = _shots(n=1, nstep=203, seed=5)[0]
    _say("graph_steps 1 / 8 / 32: bitwise identical")
    base, _ = simulate(cfg, options={"graph_steps": 8})
    for g in (1, 32):
        s, _ = simulate(cfg, options={"graph_steps": g})
        assert np.array_equal(s, base)


@pytest.mark.parametrize("wpe", [6, 8])
def test_occupancy_variant_changes_nothing(wpe):
    cfg = _shots(n=1, nstep=203, seed=5)[0]
    _say(f"waves_per_eu={wpe}: bitwise identical to the default, at creation and "
         f"switched between runs")
    base, _ = simulate(cfg)
    s, _ = simulate(cfg, options={"waves_per_eu": wpe, "graph_steps": 32})
    assert np.array_equal(s, base)

    # two runs of 100 + 103 steps (the split itself rounds differently from
    # one run: the first run ends with a separate corrector, the second
    # starts with a separate predictor), the layout switched in between
    def split(switch):
        ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                      nsbs=cfg.nsbs, seismo_types=cfg.seismo_types)
        ctx.initialize()
        ctx.run_forward(0, 100)
        if switch:
            ctx.set_launch(32, wpe)
        ctx.run_forward(100, 103)
        ctx.synchronize()
        return ctx.seismograms()

    ref = split(False)
    assert rel_l2(ref, base) <= 1e-6
    assert np.array_equal(split(True), ref)


def test_set_launch_rejects_bad_values():
    cfg = _shots(n=1, nstep=20, seed=5)[0]
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    for g, w in ((0, 0), (257, 0), (8, 5), (8, 9)):
        with pytest.raises(SmxError):
            ctx.set_launch(g, w)
    c8 = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=8)
    with pytest.raises(SmxError):
        c8.set_launch(8, 8)
    c8.set_launch(32, 0)
    with pytest.raises(SmxError):
        Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=8,
                options={"waves_per_eu": 8})


def test_autotune_times_every_candidate():
    shots = _shots(n=2, nstep=50, seed=23)
    _say("autotuner on a 2-shot batch: 3 tile sizes x 2 graph lengths x 3 "
         "occupancy variants")
    best, timings = autotune(shots, use_cache=False)
    cands = candidates(shots)
    assert len(cands) == 18
    assert sorted(timings) == sorted(label(c) for c in cands)
    assert all(np.isfinite(v) and v > 0 for v in timings.values())
    assert best in cands
    assert timings[label(best)] == min(timings.values())
    got, _ = simulate_batch(shots, options=best)
    ref, _ = simulate_batch(shots, options={"tile_elems": best["tile_elems"]})
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_batched_adjoint_equals_individual_gradients():
    shots = _shots(n=3, nstep=120, seed=19)
    _say("batched adjoint of 3 shots vs per-shot adjoints (bitwise)")
    rng = np.random.default_rng(0)
    ctx = Context.batch(shots)
    ctx.initialize()
    ctx.run_forward_checkpointed(every=50)
    gsrc = rng.standard_normal((ctx.nstep, ctx.nrec, 2)).astype(np.float32) * 1e-9
    ctx.run_adjoint(gsrc)
    g = ctx.gradient()
    for k, cfg in enumerate(shots):
        c = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
        c.initialize()
        c.run_forward_checkpointed(every=50)
        r0, r1 = ctx.shot_rec_off[k], ctx.shot_rec_off[k + 1]
        c.run_adjoint(gsrc[:, r0:r1])
        e0, e1 = ctx.shot_elem_off[k], ctx.shot_elem_off[k + 1]
        assert np.array_equal(g[:, e0:e1], c.gradient()), cfg.name


def test_batch_rejects_mismatched_shots():
    a, b = _shots(n=2, nstep=50)
    b.nstep = 60
    with pytest.raises(ValueError):
        Context.batch([a, b])
