"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
reference's SPECFEM2D golden traces.

Tolerances: seismograms <= 1e-5 relative L2 against the FP32 oracle on the
same inputs (FP32 arithmetic; the two differ only in rounding order), and the
reference's own criterion (per-station L1 ratio <= 1e-3) against the
SPECFEM2D traces.
"""

import sys

import numpy as np
import pytest

from specfempp_amd import _lib
from specfempp_amd.solver import Context, TimeMarching, simulate
from tests._cases import (FIXTURES, fixture_case, reference_metric, rel_l2, small_homogeneous,
                          small_layered)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_run(oracle_lib, cfg, precision=4, nstep=None):
    n = cfg.nstep if nstep is None else nstep
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, n, precision=precision)
    return o.run(nsbs=cfg.nsbs, types=cfg.seismo_types), o


def test_native_library_is_in_tree():
    lib = _lib.load()
    assert lib._name == _lib.LIB_PATH
    assert lib.smx_version().decode().startswith("specfempp_amd")


def test_mass_matrix_matches_oracle(oracle_lib):
    cfg = small_homogeneous(nstep=10)
    _say("mass matrix: homogeneous 80x60")
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    minv = ctx.get_field("minv").astype(np.float64)
    ib = ctx.ibool().reshape(-1)
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ominv = o.get("minv").reshape(-1, 2)
    oib = o.get("ibool").astype(np.int64)
    # compare per local point through each side's own numbering
    np.testing.assert_allclose(minv[ib], ominv[oib], rtol=2e-6)
    assert ctx.nglob == o.nglob == 77361


@pytest.mark.parametrize("fused", [True, False])
def test_forward_matches_oracle_homogeneous(oracle_lib, fused):
    cfg = small_homogeneous(nstep=800)
    _say(f"forward homogeneous 80x60, 800 steps, fused={fused}")
    ours, ctx = simulate(cfg, fused=fused)
    ref, _ = _oracle_run(oracle_lib, cfg)
    assert ours.shape == ref.shape
    err = rel_l2(ours, ref)
    _say(f"  rel L2 vs oracle = {err:.3e}")
    assert np.isfinite(ours).all()
    assert err <= 1e-5


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_specfem2d_fixtures(oracle_lib, name):
    """Displacement seismograms within 1e-5 of the FP32 oracle.  Velocity and
    acceleration seismograms carry an FP32 rounding floor of their own (the
    acceleration is M^-1 K u, a cancellation-heavy difference): for them the
    bound is twice the oracle's own FP32-vs-FP64 distance, and never looser
    than 1e-5 when that floor is lower."""
    cfg, g = fixture_case(name)
    _say(f"fixture {name}")
    ours, ctx = simulate(cfg)
    ref, _ = _oracle_run(oracle_lib, cfg)
    ref64, _ = _oracle_run(oracle_lib, cfg, precision=8)
    metric = reference_metric(ours, g["traces"])
    _say(f"  reference metric vs SPECFEM2D = {metric:.3e}")
    for it, t in enumerate(cfg.seismo_types):
        err = rel_l2(ours[:, it], ref[:, it])
        floor = rel_l2(ref[:, it], ref64[:, it])
        _say(f"  {t}: rel L2 vs oracle = {err:.3e} (oracle fp32-vs-fp64 floor {floor:.3e})")
        if t == "displacement":
            assert err <= 1e-5
        else:
            assert err <= max(1e-5, 2.0 * floor)
    assert metric < 1e-3


@pytest.mark.parametrize("fused", [True, False])
def test_sampling_every_third_step_with_pressure(oracle_lib, fused):
    """nstep_between_samples = 3 over 301 steps (not a multiple of 3) with all
    four elastic types (displacement, velocity, acceleration and the elastic
    pressure), on the fused and the operator-by-operator path, against the
    FP32 oracle sampled the same way (sample k = the state after step 3k):
    displacement and pressure within 1e-5, velocity and acceleration within
    twice the oracle's own FP32-vs-FP64 floor (at least 1e-5)."""
    import dataclasses
    cfg0, _ = fixture_case("HomogeneousElasticDomain")
    cfg = dataclasses.replace(cfg0, nstep=301, nsbs=3)
    _say(f"nsbs=3, 301 steps, d/v/a/p, fused={fused}")
    ours, ctx = simulate(cfg, fused=fused)
    ref, _ = _oracle_run(oracle_lib, cfg)
    ref64, _ = _oracle_run(oracle_lib, cfg, precision=8)
    assert ours.shape == ref.shape == (101,) + ref.shape[1:]
    assert np.all(ours[:, cfg.seismo_types.index("pressure"), :, 1] == 0)
    for it, t in enumerate(cfg.seismo_types):
        err = rel_l2(ours[:, it], ref[:, it])
        floor = rel_l2(ref[:, it], ref64[:, it])
        _say(f"  {t}: rel L2 vs oracle = {err:.3e} (floor {floor:.3e})")
        bound = 1e-5 if t in ("displacement", "pressure") else max(1e-5, 2.0 * floor)
        assert err <= bound, t


def test_layered_matches_oracle(oracle_lib):
    cfg = small_layered(nstep=400)
    _say("forward layered 120x90, 400 steps")
    ours, _ = simulate(cfg)
    ref, _ = _oracle_run(oracle_lib, cfg)
    err = rel_l2(ours, ref)
    _say(f"  rel L2 vs oracle = {err:.3e}")
    assert err <= 1e-5


def test_fp64_twin(oracle_lib):
    """The FP64 instantiation of the same kernels: close to the FP64 oracle,
    and the FP32 path's distance from it is the stated FP32 error."""
    cfg = small_homogeneous(nstep=800)
    _say("fp64 twin, homogeneous 80x60, 800 steps")
    s64, _ = simulate(cfg, precision=8)
    s32, _ = simulate(cfg, precision=4)
    ref64, _ = _oracle_run(oracle_lib, cfg, precision=8)
    e64 = rel_l2(s64, ref64)
    e32 = rel_l2(s32, s64)
    _say(f"  fp64 GPU vs fp64 oracle = {e64:.3e}; fp32 GPU vs fp64 GPU = {e32:.3e}")
    # the FP64 GPU run is fed the FP32-rounded geometry and time function
    # (it forms the moduli and impedances in FP64 from the FP32 model); the
    # FP64 oracle assembles everything in FP64 -> ~1e-7 input difference
    assert e64 <= 1e-5
    assert e32 <= 1e-5


def test_deterministic_and_fused_matches_operator_path():
    """The coloured scatter has a fixed summation order: two runs are
    bitwise identical, and the fused step matches the operator-by-operator
    loop."""
    cfg = small_homogeneous(nstep=300)
    _say("determinism, homogeneous 80x60, 300 steps")
    a, _ = simulate(cfg, fused=True)
    b, _ = simulate(cfg, fused=True)
    c, _ = simulate(cfg, fused=False)
    assert np.array_equal(a, b)
    err = rel_l2(c, a)
    _say(f"  fused vs operator path rel L2 = {err:.3e}")
    assert err <= 1e-6


def test_operator_errors():
    cfg = small_homogeneous(nstep=4)
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    n = __import__("ctypes").c_int()
    # update before initialize is a state error
    assert ctx.lib.smx_update_wavefields(ctx.h, 0, 0, __import__("ctypes").byref(n)) == -3
    ctx.initialize()
    assert ctx.lib.smx_update_wavefields(ctx.h, 0, 99, __import__("ctypes").byref(n)) == -1
    assert ctx.lib.smx_run_forward(ctx.h, 0, 5) == -1
    assert b"nstep" in ctx.lib.smx_last_error()
    with pytest.raises(_lib.SmxError):
        ctx.run_forward(2, 3)
    tm = TimeMarching(ctx)
    tm.run(fused=False)
    assert np.isfinite(ctx.seismograms()).all()


def test_geometry_from_nodes_is_bitwise(oracle_lib):
    """Q4 meshes re-form the per-point Jacobian on the device from the control
    nodes; the set-up check must find every stored FP32 value reproduced bit
    for bit, and the run keeps the oracle parity."""
    import ctypes
    cfg = small_layered(nstep=200)
    _say("geometry from control nodes: layered 120x90, 200 steps")
    ours, ctx = simulate(cfg, options={"geometry_from_nodes": True})
    gm, mis = ctypes.c_int(), ctypes.c_longlong()
    _lib.check(ctx.lib.smx_ctx_geometry(ctx.h, ctypes.byref(gm), ctypes.byref(mis)))
    _say(f"  geometry mode {gm.value}, mismatches {mis.value}")
    assert mis.value == 0 and gm.value == 2
    ref, _ = _oracle_run(oracle_lib, cfg)
    assert rel_l2(ours, ref) <= 1e-5
    stored, ctx2 = simulate(cfg)
    _lib.check(ctx2.lib.smx_ctx_geometry(ctx2.h, ctypes.byref(gm), ctypes.byref(mis)))
    assert gm.value != 2
    # same geometry bits, same arithmetic after it
    assert rel_l2(ours, stored) <= 1e-6


@pytest.mark.parametrize("name", ["HomogeneousElasticDomain",
                                  "HomogeneousElasticDomainStaceyDirichletBC"])
def test_database_mesh_fixtures(oracle_lib, name):
    """The reference's own database.bin of the fixture (read by
    specfempp_amd.meshdb) through the fused GPU path: against the oracle on
    the same mesh (1e-5) and against the SPECFEM2D traces with the
    reference's criterion."""
    import dataclasses
    import os
    from specfempp_amd.meshdb import read_mesh_database_2d
    from tests._cases import GOLDEN, fixture_case, reference_metric
    cfg, g = fixture_case(name)
    mesh, _ = read_mesh_database_2d(os.path.join(GOLDEN, f"specfem2d_{name}.database.bin.gz"))
    cfg = dataclasses.replace(cfg, mesh=mesh)
    _say(f"database mesh {name}: GPU vs oracle and SPECFEM2D")
    ours, _ = simulate(cfg)
    ref = oracle_lib.Oracle2D(mesh, cfg.sources, cfg.receivers, cfg.dt,
                              cfg.nstep).run(types=cfg.seismo_types)
    e_or, e_sp = rel_l2(ours[:, 0], ref[:, 0]), reference_metric(ours, g["traces"])
    _say(f"  vs oracle {e_or:.3e}; vs SPECFEM2D (reference metric) {e_sp:.3e}")
    assert e_or <= 1e-5
    assert e_sp < 1e-3
