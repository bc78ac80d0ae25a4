"""GPU parity on curved and skewed meshes (CPU oracle as the checker): the
reference's fixtures are all rectangular, so these cases exercise what they
leave open -- per-point geometry that is not element-constant, Stacey
normals and free surfaces on curved edges, coupling normals on a curved
fluid-solid interface, and point location inside curved elements.

Each case: a box mesh whose control nodes are moved smoothly (interior
sine pattern up to 0.18 element sizes, sinusoidal topography of 0.25
element sizes on top), heterogeneous media, sources and stations away from
the boundary.  Bound per seismogram type: twice the type's FP32 floor
(oracle FP32 vs FP64), at least 1e-5, against both oracles.
"""

import sys

import numpy as np
import pytest

from specfempp_amd.solver import simulate
from tests._cases import acoustic_box, adjoint_small, coupled_box, rel_l2, warped

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(oracle_lib, cfg):
    ours, _ = simulate(cfg)
    o32 = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    o64 = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                              precision=8)
    r32, r64 = o32.run(types=cfg.seismo_types), o64.run(types=cfg.seismo_types)
    bad = []
    for it, t in enumerate(cfg.seismo_types):
        e, e64 = rel_l2(ours[:, it], r32[:, it]), rel_l2(ours[:, it], r64[:, it])
        bound = max(1e-5, 2.0 * rel_l2(r32[:, it], r64[:, it]))
        _say(f"  {t}: rel L2 vs FP32 oracle {e:.3e}, vs FP64 oracle {e64:.3e} "
             f"(bound {bound:.1e})")
        assert np.isfinite(ours[:, it]).all() and np.abs(ours[:, it]).max() > 0
        if not (e <= bound and e64 <= bound):
            bad.append((t, e, e64, bound))
    assert not bad, bad


@pytest.mark.parametrize("ngnod", [4, 9])
def test_warped_elastic(oracle_lib, ngnod):
    from specfempp_amd.mesh import box_mesh
    import dataclasses
    cfg = adjoint_small(nx=16, nz=12, nstep=400)
    if ngnod == 9:
        m9 = box_mesh(0.0, 16 * 50.0, 0.0, 12 * 50.0, 16, 12, ngnod=9,
                      absorbing=("bottom", "right", "left"))
        cfg = dataclasses.replace(cfg, mesh=dataclasses.replace(
            m9, kappa=cfg.mesh.kappa, mu=cfg.mesh.mu, rho=cfg.mesh.rho))
    cfg = dataclasses.replace(warped(cfg), seismo_types=("displacement", "velocity"))
    _say(f"warped elastic box 16x12 (ngnod {ngnod}), Stacey on a curved mesh, 400 steps")
    _check(oracle_lib, cfg)


def test_warped_acoustic(oracle_lib):
    cfg = warped(acoustic_box(nx=16, nz=12, nstep=300, ngnod=9))
    _say("warped acoustic box 16x12 (Q9), curved free surface, 300 steps, d/v/a/p")
    _check(oracle_lib, cfg)


def test_warped_coupled(oracle_lib):
    cfg = warped(coupled_box(nx=16, nz=12, nstep=300, vertical=False, ngnod=9))
    _say("warped coupled box 16x12 (Q9), curved fluid-solid interface, 300 steps")
    _check(oracle_lib, cfg)
