"""GPU parity of coupled elastic-acoustic 2D meshes (through the C ABI)
against the coupled oracle and the reference's two coupled SPECFEM2D
fixtures.

Cases: small heterogeneous boxes with a vertical (water left) and a
horizontal (water on top) interface, free water surface, Stacey on three
sides, a source in each medium, stations across the interface (d, v, a);
the fixture with the water above the solid (restated mesh) and the one with
the water below (the reference mesher's database).

Tolerances: per seismogram type twice the type's FP32 floor (oracle FP32 vs
FP64, measured in the test), at least 1e-5, against the FP32 and the FP64
oracle; against SPECFEM2D the oracle's own bound (reference metric 1e-4);
the medium-tagged operator path against the fused path and the oracle the
same FP32-floor bound (the two paths round the coupled accelerations in a
different order: 6e-7 in displacement, 2e-6 in velocity).
"""

import sys

import numpy as np
import pytest

from specfempp_amd.solver import Context, simulate
from tests._cases import (coupled_box, coupled_fixture_case, coupled_variant2_case,
                          reference_metric, rel_l2)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle(oracle_lib, cfg, precision):
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=precision)
    assert o.coupled
    return o.run(nsbs=cfg.nsbs, types=cfg.seismo_types)


@pytest.mark.parametrize("vertical", [True, False])
def test_coupled_box_matches_oracle(oracle_lib, vertical):
    cfg = coupled_box(vertical=vertical)
    _say(f"coupled box ({'vertical' if vertical else 'horizontal'} interface), 400 steps")
    ours, ctx = simulate(cfg)
    assert ctx.ncomp == 0
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    bad = []
    for it, t in enumerate(cfg.seismo_types):
        e, e64 = rel_l2(ours[:, it], ref32[:, it]), rel_l2(ours[:, it], ref64[:, it])
        bound = max(1e-5, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
        _say(f"  {t}: rel L2 vs FP32 oracle {e:.3e}, vs FP64 oracle {e64:.3e} "
             f"(bound {bound:.1e})")
        assert np.isfinite(ours[:, it]).all()
        if not (e <= bound and e64 <= bound):
            bad.append((t, e, e64, bound))
    assert not bad, bad


def test_coupled_fixtures_match_specfem2d(oracle_lib):
    for make in (coupled_fixture_case, coupled_variant2_case):
        cfg, traces = make()
        _say(f"coupled fixture {cfg.name}: 144x108 Q9, 600 steps")
        ours, ctx = simulate(cfg)
        np.testing.assert_allclose(ctx.t0, -0.12, atol=1e-6)
        e = reference_metric(ours, traces)
        _say(f"  reference metric vs SPECFEM2D {e:.3e}")
        assert e < 1e-4
    ref = _oracle(oracle_lib, cfg, 4)
    e = rel_l2(ours, ref)
    _say(f"  {cfg.name}: rel L2 vs FP32 oracle {e:.3e}")
    assert e < 1e-5


def test_coupled_run_is_reproducible_and_tiles_agree(oracle_lib):
    cfg = coupled_box(nstep=200)
    _say("coupled box: reruns bitwise equal; 16-element tiles within the FP32 floor")
    a, _ = simulate(cfg)
    b, _ = simulate(cfg)
    assert np.array_equal(a, b)
    c, _ = simulate(cfg, options={"tile_elems": 16, "graph_steps": 3})
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    for it in range(len(cfg.seismo_types)):
        bound = max(1e-6, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
        assert rel_l2(c[:, it], a[:, it]) <= bound


@pytest.mark.parametrize("vertical", [True, False])
def test_coupled_operator_path(oracle_lib, vertical):
    """The reference loop's medium-tagged operator sequence on a coupled mesh
    (predictors, update<acoustic>, corrector(acoustic), update<elastic_psv>,
    corrector(elastic_psv), seismograms of both media) against the fused
    path and the oracle, with the loop's DOF and element checks."""
    cfg = coupled_box(vertical=vertical, nstep=300)
    _say(f"coupled operator path ({'vertical' if vertical else 'horizontal'}), 300 steps")
    fused, _ = simulate(cfg)
    ops, ctx = simulate(cfg, fused=False)
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    for it, t in enumerate(cfg.seismo_types):
        e_f = rel_l2(ops[:, it], fused[:, it])
        e_o = rel_l2(ops[:, it], ref32[:, it])
        bound = max(1e-5, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
        _say(f"  {t}: operator vs fused {e_f:.3e}; vs FP32 oracle {e_o:.3e} (bound {bound:.1e})")
        assert e_f <= bound and e_o <= bound


def test_coupled_medium_counts():
    """Each medium-tagged operator counts its own medium: the solid's DOFs
    are 2 per point, the fluid's 1, unknown media count 0, and together they
    cover the context's degrees of freedom and elements."""
    import ctypes
    from specfempp_amd import _lib
    cfg = coupled_box(nstep=10)
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    n = ctypes.c_int()
    counts = {}
    for name, tag in _lib.MEDIUM_TAGS.items():
        _lib.check(ctx.lib.smx_predictor_medium(ctx.h, 0, tag, ctypes.byref(n)))
        d = n.value
        _lib.check(ctx.lib.smx_update_wavefields_medium(ctx.h, 0, tag, 0, ctypes.byref(n)))
        counts[name] = (d, n.value)
    total = ctypes.c_longlong()
    _lib.check(ctx.lib.smx_ctx_degrees_of_freedom(ctx.h, ctypes.byref(total)))
    acoustic = int(cfg.mesh.acoustic.sum())
    assert counts["acoustic"][1] == acoustic
    assert counts["elastic_psv"][1] == cfg.mesh.nspec - acoustic
    assert counts["elastic_psv"][0] % 2 == 0
    assert counts["acoustic"][0] + counts["elastic_psv"][0] == total.value
    assert all(v == (0, 0) for k, v in counts.items() if k not in ("acoustic", "elastic_psv"))
    # the adjoint field is not held on coupled meshes
    assert ctx.lib.smx_predictor_medium(ctx.h, 1, 0, ctypes.byref(n)) == _lib.SMX_ERR_ARG


def test_coupled_context_limits():
    cfg = coupled_box(nstep=10)
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    with pytest.raises(RuntimeError, match="one field per medium"):
        ctx.get_field("u")
    import ctypes
    n = ctypes.c_int()
    assert ctx.lib.smx_predictor(ctx.h, 0, ctypes.byref(n)) == -3
    assert b"smx_run_forward" in ctx.lib.smx_last_error()
    with pytest.raises(RuntimeError, match="FP32"):
        Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=8)
