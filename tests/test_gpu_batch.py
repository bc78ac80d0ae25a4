"""Batched shots and layout options on the GPU.

* A batched context advances several independent shots with one launch per
  kernel per step; every shot keeps its own tiling and summation order, so
  its seismograms (and adjoint gradients) equal those of its own context bit
  for bit.
* Tile sizes 16 / 32 / 64 only change summation order: seismograms stay
  within 1e-5 of the oracle.  The graph length changes nothing.
"""

import sys

import numpy as np
import pytest

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
