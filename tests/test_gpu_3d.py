"""3D GPU path (through the C ABI) against the 3D oracle and the reference's
SPECFEM3D golden traces.

Tolerances: seismograms <= 1e-5 relative L2 against the FP32 oracle on the
same inputs; the reference's criterion (per-station L1 ratio <= 1e-3)
against SPECFEM3D.
"""

import sys

import numpy as np
import pytest

from specfempp_amd.mesh3d import (Config3D, Receivers3D, Sources3D, box_mesh3d,
                                  halfspace_fixture)
from specfempp_amd.solver3d import Context3D, simulate3d
from tests._cases import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _metric(ours, traces):
    worst = 0.0
    for it in range(traces.shape[0]):
        for ista in range(traces.shape[1]):
            a = np.asarray(ours[:, it, ista, :], dtype=np.float64).T
            b = traces[it, ista]
            worst = max(worst, float(np.abs(a - b).sum() / np.abs(b).sum()))
    return worst


def test_3d_halfspace_fixture(oracle_lib):
    g = np.load(f"{GOLDEN}/specfem3d_HomogeneousHalfspaceSmallNoABCForceSource.npz")
    mesh, src = halfspace_fixture()
    rec = Receivers3D(x=g["x"], y=g["y"], z=g["z"])
    types = ("displacement", "velocity", "acceleration")
    cfg = Config3D("halfspace", mesh, src, rec, 0.035, 75, seismo_types=types)
    _say("3D halfspace fixture (8x6x6 hex), 75 steps, d/v/a")
    s, ctx = simulate3d(cfg)
    assert ctx.nglob == 33 * 25 * 25
    _say(f"  geometry {ctx.geometry}")
    err_ref = _metric(s, g["traces"])
    ref = oracle_lib.Oracle3D(mesh, src, rec, 0.035, 75).run(types=types)
    errs = [rel_l2(s[:, k], ref[:, k]) for k in range(3)]
    _say(f"  vs SPECFEM3D (reference metric) {err_ref:.3e}; vs oracle d/v/a {errs}")
    assert err_ref <= 1e-3
    assert errs[0] <= 1e-5
    assert max(errs) <= 5e-5  # v, a: FP32 cancellation floor as in 2D


def _stacey_case(nstep=300, moment=True):
    mesh = box_mesh3d(0.0, 1200.0, 0.0, 1000.0, -800.0, 0.0, 6, 5, 4,
                      rho=lambda x, y, z: 2500.0 + 0.1 * (x + y),
                      vp=lambda x, y, z: 3000.0 - 0.5 * z, vs=lambda x, y, z: 1700.0 - 0.3 * z)
    if moment:
        src = Sources3D.make([610.0], [480.0], [-390.0],
                             moment=(1e13, -5e12, 2e12, 3e12, 1e12, -2e12), f0=12.0)
    else:
        src = Sources3D.make([610.0], [480.0], [-390.0], force=(1e10, 0.0, -2e10), f0=12.0)
    rx = np.array([100.0, 400.0, 700.0, 1100.0, 600.0])
    ry = np.array([100.0, 900.0, 500.0, 300.0, 500.0])
    rz = np.array([0.0, 0.0, -100.0, -700.0, -400.0])
    rec = Receivers3D(x=rx, y=ry, z=rz)
    return Config3D("stacey3d", mesh, src, rec, 1e-3, nstep)


@pytest.mark.parametrize("moment", [True, False])
def test_3d_stacey_heterogeneous_matches_oracle(oracle_lib, moment):
    cfg = _stacey_case(moment=moment)
    _say(f"3D heterogeneous box 6x5x4, Stacey on 5 faces, moment={moment}, 300 steps")
    s, _ = simulate3d(cfg)
    ref = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep).run()
    err = rel_l2(s, ref)
    s64, _ = simulate3d(cfg, precision=8)
    ref64 = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                                precision=8).run()
    e64 = rel_l2(s64, ref64)
    e32 = rel_l2(s, s64)
    _say(f"  fp32 vs oracle {err:.3e}; fp64 vs fp64 oracle {e64:.3e}; fp32 vs fp64 {e32:.3e}; "
         f"geometry {_.geometry}")
    assert err <= 1e-5
    assert e64 <= 1e-5
    assert e32 <= 1e-5


def test_3d_warped_mesh_matches_oracle(oracle_lib):
    """Curved hex8 elements (interior warp, topography): per-point geometry
    that is not element-constant, Stacey on curved faces, the moment-tensor
    source gradient and point location inside curved elements."""
    from tests._cases import warped3d
    cfg = warped3d(_stacey_case(moment=True))
    _say("3D warped box 6x5x4, Stacey on 5 curved faces, moment tensor, 300 steps")
    s, ctx = simulate3d(cfg)
    assert ctx.layout["geom_elem_elements"] == 0
    ref = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep).run()
    ref64 = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                                precision=8).run()
    err, e64, floor = rel_l2(s, ref), rel_l2(s, ref64), rel_l2(ref, ref64)
    _say(f"  fp32 vs oracle {err:.3e}; vs fp64 oracle {e64:.3e}; oracle floor {floor:.3e}")
    bound = max(1e-5, 2.0 * floor)
    assert err <= bound and e64 <= bound


def test_3d_hex_workload_generator_at_12cubed_matches_oracle(oracle_lib):
    """The 64^3 workload's generator at 12^3 (same 6.4 km cube, random
    medium, explosion source at 3 km depth, 256 surface receivers, free top,
    Stacey on five faces), dt scaled with the element size, 300 steps (2 s):
    FP32 against the FP32 oracle and against the FP64 run of the same code."""
    from specfempp_amd.mesh3d import hex_box_case
    h = 6400.0 / 12
    cfg = hex_box_case(n=12, h=h, dt=1.25e-3 * h / 100.0, nstep=300)
    _say("3D workload generator at 12^3, 300 steps")
    s, ctx = simulate3d(cfg)
    ref = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep).run()
    s64, _ = simulate3d(cfg, precision=8)
    err, e32 = rel_l2(s, ref), rel_l2(s, s64)
    _say(f"  fp32 vs oracle {err:.3e}; fp32 vs fp64 {e32:.3e}; geometry {ctx.geometry}")
    assert np.abs(ref).max() > 0
    assert err <= 1e-5
    assert e32 <= 1e-5


def _kinetic(ctx):
    v = ctx.get_field("v").astype(np.float64)
    m = 1.0 / ctx.get_field("minv").astype(np.float64)
    return float(np.sum(m * v * v))


@pytest.mark.parametrize("absorbing", [True, False])
def test_3d_absorbing_boundary_removes_energy(absorbing):
    """Kinetic energy sum(M v^2) at step 1500 against step 300: with Stacey on
    five faces (free top) the oracle gives 6.6 %; with every face free the
    energy stays in the box (97 %)."""
    cfg = _stacey_case(nstep=1500)
    if not absorbing:
        cfg.mesh.abs_ispec = cfg.mesh.abs_ispec[:0]
        cfg.mesh.abs_face = cfg.mesh.abs_face[:0]
    _say(f"3D kinetic energy decay, absorbing={absorbing}")
    ctx = Context3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    ctx.run_forward(0, 300)
    ctx.synchronize()
    e_early = _kinetic(ctx)
    ctx.run_forward(300, 1200)
    ctx.synchronize()
    e_late = _kinetic(ctx)
    _say(f"  E_kin step 300: {e_early:.3e}, step 1500: {e_late:.3e}")
    assert np.isfinite(e_late)
    if absorbing:
        assert e_late < 0.1 * e_early
    else:
        assert e_late > 0.5 * e_early


def test_3d_errors():
    cfg = _stacey_case(nstep=10)
    ctx = Context3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    with pytest.raises(RuntimeError, match="initialize"):
        ctx.run_forward(0, 5)
    ctx.initialize()
    with pytest.raises(RuntimeError, match="steps outside"):
        ctx.run_forward(5, 10)
    from specfempp_amd import _lib
    import ctypes
    n = ctypes.c_int()
    assert _lib.load().smx_update_wavefields(ctx.h, 0, 0, ctypes.byref(n)) == _lib.SMX_ERR_STATE


def test_3d_geometry_from_nodes_is_exact():
    """Coordinates far from the origin leave FP32 rounding in the per-point
    Jacobian (the elements are not flagged affine); the device re-forms the
    stored values from the control nodes bit for bit, and the run matches a
    context that reads the stored values."""
    mesh = box_mesh3d(37000.0, 37600.0, 52000.0, 52500.0, -3400.0, -3000.0, 6, 5, 4,
                      rho=2600.0, vp=3200.0, vs=1800.0)
    src = Sources3D.make([37310.0], [52240.0], [-3210.0], force=(0.0, 1e10, 1e10), f0=15.0)
    rec = Receivers3D(x=np.array([37100.0, 37500.0]), y=np.array([52100.0, 52400.0]),
                      z=np.array([-3000.0, -3300.0]))
    cfg = Config3D("far", mesh, src, rec, 5e-4, 200)
    s, ctx = simulate3d(cfg, options={"geometry_from_nodes": True})
    _say(f"3D far-from-origin box: geometry {ctx.geometry}")
    assert ctx.geometry["mode"] == "nodes" and ctx.geometry["mismatches"] == 0
    s_stored, ctx2 = simulate3d(cfg)
    assert ctx2.geometry["mode"] == "point"
    assert np.linalg.norm(s - s_stored) <= 1e-6 * np.linalg.norm(s_stored)
