"""The two seismogram paths of the fused 2D elastic step give bitwise the same
displacement traces.

A run that records displacement (and pressure) only forms every receiver
term in the tile pass, from the displacement each tile gathered, and sums the
25 terms per receiver in the shared-point pass's first workgroup.  A run that
also records velocity takes the terms on shared points from the displacement
field in a separate seismogram pass.  Both sum the same rounded products in
the same point order, so the displacement channel must agree bit for bit:
every step and every third step, a batch of shots, and FP64.
"""

import dataclasses
import sys

import numpy as np
import pytest

from specfempp_amd.config import batched_shots, homogeneous_box
from specfempp_amd.solver import Context

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _forward(cfg, types, nsbs=1, precision=4):
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, nsbs=nsbs,
                  seismo_types=types, precision=precision)
    ctx.initialize()
    ctx.run_forward(0, cfg.nstep)
    ctx.synchronize()
    s = ctx.seismograms()
    ctx.close()
    return s


@pytest.mark.parametrize("nsbs,precision", [(1, 4), (3, 4), (1, 8)])
def test_displacement_paths_bitwise(nsbs, precision):
    cfg = homogeneous_box(nstep=601)
    _say(f"seismogram paths: 80x60 box, 601 steps, nsbs={nsbs}, precision={precision}")
    d = _forward(cfg, ("displacement",), nsbs, precision)
    dv = _forward(cfg, ("displacement", "velocity"), nsbs, precision)
    pd = _forward(cfg, ("pressure", "displacement"), nsbs, precision)
    assert np.isfinite(d).all() and np.abs(d).max() > 0
    np.testing.assert_array_equal(d[:, 0], dv[:, 0])
    np.testing.assert_array_equal(d[:, 0], pd[:, 1])


def test_displacement_paths_bitwise_batched():
    shots = batched_shots(nshots=3, nstep=400, nmax=120)
    _say("seismogram paths: 3 batched shots, 400 steps")
    out = {}
    for types in [("displacement",), ("displacement", "acceleration")]:
        batch = [dataclasses.replace(c, seismo_types=types) for c in shots]
        ctx = Context.batch(batch)
        ctx.initialize()
        ctx.run_forward(0, 400)
        ctx.synchronize()
        out[types] = ctx.seismograms()
        ctx.close()
    d, da = out[("displacement",)], out[("displacement", "acceleration")]
    assert np.isfinite(d).all() and np.abs(d).max() > 0
    np.testing.assert_array_equal(d[:, 0], da[:, 0])
