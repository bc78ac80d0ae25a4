"""GPU parity of the 2D acoustic path (through the C ABI) against the CPU
oracle and the reference's acoustic SPECFEM2D golden traces.

Cases: the three acoustic fixtures (80x60 Q9 box from the reference mesher's
databases; Dirac source with free surfaces on all sides; four Ricker sources
with free surfaces on all sides; the same with Stacey on three sides and a
free top) and a heterogeneous acoustic box (smooth +-5% rho and vp, free top,
Stacey on three sides, two Ricker sources and one Dirac, d/v/a/p).

Tolerances (relative L2 per seismogram type):
* FP32 GPU vs FP32 oracle and vs FP64 oracle: twice the FP32 floor of the
  type (oracle FP32 vs oracle FP64, measured in the test), at least 1e-5.
  Every acoustic type is a cancellation: d / v / a are the gradient of the
  potential (rho^-1 grad chi), pressure is M^-1 (f - K chi - B chi_dot); on
  the 80x60 fixtures the FP32 floor of the displacement is 1.5e-5.
* FP64 GPU vs FP64 oracle: the larger of 1e-5 and the type's FP32 floor.
  As for the elastic FP64 twin, the FP64 GPU run is fed the FP32-rounded
  geometry, time function and receiver interpolants of the host assembly (it
  forms 1/rho and 1/(rho vp) in double from the FP32 model) while the FP64
  oracle assembles everything in double: a ~1e-7 input difference, which
  each type amplifies as it amplifies FP32 rounding.
* Against SPECFEM2D: the reference's criterion with the same bounds as the
  oracle's own test.
"""

import dataclasses
import sys

import numpy as np
import pytest

from specfempp_amd.solver import Context, simulate
from tests._cases import (ACOUSTIC_FIXTURES, acoustic_box, acoustic_fixture_case,
                          reference_metric, rel_l2)

pytestmark = pytest.mark.gpu

GOLDEN_BOUNDS = {"HomogeneousAcousticDomain": 5e-4,
                 "HomogeneousAcousticDomainDirichletBC": 1e-4,
                 "HomogeneousAcousticDomainCompositeStaceyDirichletBC": 1e-4}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle(oracle_lib, cfg, precision):
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=precision)
    return o.run(nsbs=cfg.nsbs, types=cfg.seismo_types)


def _check_types(cfg, ours, ref32, ref64):
    bad = []
    for it, t in enumerate(cfg.seismo_types):
        e = rel_l2(ours[:, it], ref32[:, it])
        e64 = rel_l2(ours[:, it], ref64[:, it])
        floor = rel_l2(ref32[:, it], ref64[:, it])
        bound = max(1e-5, 2.0 * floor)
        _say(f"  {t}: rel L2 vs FP32 oracle {e:.3e}, vs FP64 oracle {e64:.3e} "
             f"(oracle FP32 floor {floor:.3e}, bound {bound:.1e})")
        assert np.isfinite(ours[:, it]).all()
        if not (e <= bound and e64 <= bound):
            bad.append((t, e, e64, bound))
    assert not bad, bad


@pytest.mark.parametrize("name", sorted(ACOUSTIC_FIXTURES))
def test_acoustic_fixtures_match_oracle_and_specfem2d(oracle_lib, name):
    cfg, traces = acoustic_fixture_case(name)
    _say(f"acoustic fixture {name}: 80x60 Q9, 300 steps, types {cfg.seismo_types}")
    ours, ctx = simulate(cfg)
    assert ctx.ncomp == 1
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    _check_types(cfg, ours, ref32, ref64)
    for it, t in enumerate(cfg.seismo_types):
        e = reference_metric(ours[:, it:it + 1], traces[it:it + 1])
        _say(f"  {t}: reference metric vs SPECFEM2D {e:.3e}")
        assert e < GOLDEN_BOUNDS[name]


@pytest.mark.parametrize("ngnod", [4, 9])
def test_acoustic_box_matches_oracle(oracle_lib, ngnod):
    cfg = acoustic_box(nstep=400, ngnod=ngnod)
    _say(f"acoustic heterogeneous box 24x18 (ngnod {ngnod}), 400 steps, d/v/a/p")
    ours, _ = simulate(cfg)
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    _check_types(cfg, ours, ref32, ref64)
    ours64, _ = simulate(cfg, precision=8)
    bad = []
    for it, t in enumerate(cfg.seismo_types):
        e = rel_l2(ours64[:, it], ref64[:, it])
        bound = max(1e-5, rel_l2(ref32[:, it], ref64[:, it]))
        _say(f"  FP64 {t}: rel L2 vs FP64 oracle {e:.3e} (bound {bound:.1e})")
        if not e <= bound:
            bad.append((t, e, bound))
    assert not bad, bad


def test_acoustic_operator_path_matches_fused(oracle_lib):
    """The operator-by-operator loop and the fused graph differ only in
    rounding (where the compiler fuses a multiply-add): per type within
    twice the type's FP32 floor (oracle FP32 vs FP64), at least 1e-6, the
    rule for two FP32 evaluations of the same step."""
    cfg = acoustic_box(nstep=120)
    _say("acoustic operator-by-operator loop vs fused graph")
    fused, cf = simulate(cfg, fused=True)
    ops, co = simulate(cfg, fused=False)
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    bad = []
    for it, t in enumerate(cfg.seismo_types):
        e = rel_l2(ops[:, it], fused[:, it])
        bound = max(1e-6, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
        _say(f"  {t}: operator vs fused rel L2 {e:.3e} (bound {bound:.1e})")
        if not e <= bound:
            bad.append((t, e, bound))
    assert not bad, bad
    assert rel_l2(co.get_field("u"), cf.get_field("u")) <= 1e-6


def test_acoustic_mass_matches_oracle(oracle_lib):
    cfg, _ = acoustic_fixture_case("HomogeneousAcousticDomainCompositeStaceyDirichletBC")
    _say("acoustic mass (with the Stacey term) vs oracle")
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    minv = ctx.get_field("minv").astype(np.float64)
    assert minv.shape == (ctx.nglob,)
    ib = ctx.ibool().reshape(-1)
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ominv = o.get("minv").reshape(-1, 2)[:, 0]
    oib = o.get("ibool").astype(np.int64)
    np.testing.assert_allclose(minv[ib], ominv[oib], rtol=2e-6)


def test_acoustic_tile_sizes_and_reruns_agree(oracle_lib):
    """Reruns are bitwise equal; other tile layouts (other summation
    orders) agree within twice each type's FP32 floor."""
    cfg = acoustic_box(nstep=200)
    _say("acoustic tile layouts 64, 16, 1 agree; reruns are bitwise equal")
    base, _ = simulate(cfg)
    again, _ = simulate(cfg)
    assert np.array_equal(base, again)
    ref32, ref64 = _oracle(oracle_lib, cfg, 4), _oracle(oracle_lib, cfg, 8)
    bad = []
    for te in (16, 1):
        s, _ = simulate(cfg, options={"tile_elems": te})
        for it, t in enumerate(cfg.seismo_types):
            e = rel_l2(s[:, it], base[:, it])
            bound = max(1e-6, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
            _say(f"  tile {te} {t}: rel L2 vs tile 64 {e:.3e} (bound {bound:.1e})")
            if not e <= bound:
                bad.append((te, t, e, bound))
    assert not bad, bad


def test_acoustic_free_surface_points_stay_at_rest():
    cfg = acoustic_box(nx=10, nz=8, nstep=200, free=("top", "left"))
    _say("acoustic free-surface points stay exactly at rest")
    _, ctx = simulate(cfg, precision=8)
    u = ctx.get_field("u")
    ib = ctx.ibool()
    m = cfg.mesh
    top = ib[m.free_ispec[m.free_type == 3], 4, :].ravel()
    left = ib[m.free_ispec[m.free_type == 4], :, 0].ravel()
    assert np.abs(u).max() > 0
    assert np.all(u[top] == 0) and np.all(u[left] == 0)


def test_acoustic_adjoint_is_rejected():
    cfg = acoustic_box(nstep=20)
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    with pytest.raises(RuntimeError, match="elastic"):
        ctx.run_forward_checkpointed(10)
