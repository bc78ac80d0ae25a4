"""The 3D oracle (oracle/cpu_ref3d.cpp) pinned against the reference's
SPECFEM3D golden traces and closed forms (CPU only).

* HomogeneousHalfspaceSmallNoABCForceSource (the reference's 3D displacement
  test, tests/unit-tests/displacement_tests/Newmark/dim3/newmark_tests.cpp):
  d, v, a at 26 stations with the reference's criterion (per-station
  sum|s - ref| / sum|ref| <= 1e-3); a wrong-polarity force fails it.
* Mass: sum of M = rho * volume; rigid translation gives no force; K is
  symmetric; a uniform strain gives no interior force.
* Moment-tensor source array: zero net force and first moments equal to M.
* Stacey faces (this build's own 3D boundary): the traction summed over an
  absorbing face is -rho vp v_n area for a normal velocity and -rho vs v_t
  area for a tangential one.
"""

import numpy as np
import pytest

from specfempp_amd.mesh3d import (Receivers3D, Sources3D, box_mesh3d, gll_coordinates3d,
                                  halfspace_fixture)
from tests._cases import GOLDEN

CASE = "HomogeneousHalfspaceSmallNoABCForceSource"


def _metric(ours, traces):
    worst = 0.0
    for it in range(traces.shape[0]):
        for ista in range(traces.shape[1]):
            a = np.asarray(ours[:, it, ista, :], dtype=np.float64).T
            b = traces[it, ista]
            worst = max(worst, float(np.abs(a - b).sum() / np.abs(b).sum()))
    return worst


@pytest.fixture(scope="module")
def golden():
    return np.load(f"{GOLDEN}/specfem3d_{CASE}.npz")


def test_specfem3d_golden_traces(oracle_lib, golden):
    mesh, src = halfspace_fixture()
    rec = Receivers3D(x=golden["x"], y=golden["y"], z=golden["z"])
    o = oracle_lib.Oracle3D(mesh, src, rec, 0.035, 75)
    assert o.nglob == 33 * 25 * 25
    assert abs(o.t0 - golden["times"][0]) < 1e-6
    s = o.run(types=("displacement", "velocity", "acceleration"))
    err = _metric(s, golden["traces"])
    print("3D oracle vs SPECFEM3D traces: worst per-station L1 ratio", err)
    assert err <= 1e-3


def test_specfem3d_golden_rejects_wrong_polarity(oracle_lib, golden):
    mesh, _ = halfspace_fixture()
    src = Sources3D.make([4000.0], [3500.0], [-3750.0], force=(0.0, 0.0, -1e14), f0=1.25)
    rec = Receivers3D(x=golden["x"], y=golden["y"], z=golden["z"])
    s = oracle_lib.Oracle3D(mesh, src, rec, 0.035, 75).run()
    assert _metric(s, golden["traces"][:1]) > 1.0


def _small(absorbing=(), n=(3, 2, 2)):
    mesh = box_mesh3d(0.0, 300.0, 0.0, 200.0, -200.0, 0.0, *n, absorbing=absorbing,
                      rho=2000.0, vp=3000.0, vs=1700.0)
    src = Sources3D.make([150.0], [100.0], [-100.0], force=(0.0, 0.0, 1.0), f0=10.0)
    rec = Receivers3D(x=np.array([100.0]), y=np.array([100.0]), z=np.array([-50.0]))
    return mesh, src, rec


def test_mass_sum_is_rho_volume(oracle_lib):
    mesh, src, rec = _small()
    o = oracle_lib.Oracle3D(mesh, src, rec, 1e-3, 1, precision=8)
    M = 1.0 / o.get("minv").reshape(-1, 3)
    np.testing.assert_allclose(M[:, 0], M[:, 2])
    assert abs(M[:, 0].sum() - 2000.0 * 300 * 200 * 200) < 1e-9 * 2000.0 * 300 * 200 * 200


def test_rigid_translation_and_uniform_strain(oracle_lib):
    mesh, src, rec = _small(n=(3, 3, 3))
    o = oracle_lib.Oracle3D(mesh, src, rec, 1e-3, 1, precision=8)
    ib = o.get("ibool").astype(np.int64).reshape(mesh.nspec, 125)
    xyz = gll_coordinates3d(mesh).reshape(mesh.nspec, 125, 3)
    pts = np.zeros((o.nglob, 3))
    pts[ib.ravel()] = xyz.reshape(-1, 3)
    u = np.tile([1.0, -2.0, 0.5], (o.nglob, 1))
    a = o.apply_stiffness(u, with_stacey=False)
    scale = 2000.0 * 3000.0 ** 2 * 100.0  # ~ lambda+2mu * element size
    assert np.abs(a).max() < 1e-9 * scale
    # uniform strain: u = E x; interior nodes feel no force
    E = np.array([[1e-3, 2e-4, 0.0], [2e-4, -5e-4, 1e-4], [0.0, 1e-4, 3e-4]])
    a = o.apply_stiffness(pts @ E.T, with_stacey=False)
    inner = np.all((pts > [1e-6, 1e-6, -200 + 1e-6]) & (pts < [300 - 1e-6, 200 - 1e-6, -1e-6]),
                   axis=1)
    assert inner.sum() > 0
    assert np.abs(a[inner]).max() < 1e-9 * np.abs(a).max()


def test_stiffness_symmetric(oracle_lib):
    mesh, src, rec = _small()
    o = oracle_lib.Oracle3D(mesh, src, rec, 1e-3, 1, precision=8)
    rng = np.random.default_rng(0)
    r1, r2 = rng.standard_normal((2, o.nglob, 3))
    k1 = o.apply_stiffness(r1, with_stacey=False)
    k2 = o.apply_stiffness(r2, with_stacey=False)
    lhs, rhs = float(np.sum(r2 * k1)), float(np.sum(r1 * k2))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    assert float(np.sum(r1 * k1)) < 0  # -K is negative definite


def test_moment_tensor_source_moments(oracle_lib):
    mesh, _, rec = _small()
    M6 = (2e15, -1e15, 5e14, 3e14, -2e14, 1e14)  # Mxx Myy Mzz Mxy Mxz Myz
    src = Sources3D.make([137.0], [91.0], [-77.0], moment=M6, f0=10.0)
    o = oracle_lib.Oracle3D(mesh, src, rec, 1e-3, 1, precision=8)
    L = o.get("source_array").reshape(3, 125)
    # the element the source sits in (1 m inside a 100 m grid cell)
    e = (0 * 2 + 0) * 3 + 1  # jx = 1, jy = 0, jz = 0
    xyz = gll_coordinates3d(mesh)[e].reshape(125, 3)
    m6 = np.asarray(M6, dtype=np.float32).astype(np.float64)  # the source's FP32 tensor
    M = np.array([[m6[0], m6[3], m6[4]], [m6[3], m6[1], m6[5]], [m6[4], m6[5], m6[2]]])
    np.testing.assert_allclose(L.sum(axis=1), 0.0, atol=1e-9 * np.abs(M).max() / 100)
    np.testing.assert_allclose(L @ xyz, M, rtol=1e-9, atol=1e-9 * np.abs(M).max())


def test_stacey_face_impedance(oracle_lib):
    mesh, src, rec = _small(absorbing=("bottom",))
    o = oracle_lib.Oracle3D(mesh, src, rec, 1e-3, 1, precision=8)
    rho, vp, vs, area = 2000.0, 3000.0, 1700.0, 300.0 * 200.0
    kappa, mu, r = (np.asarray(x, np.float64) for x in (mesh.kappa, mesh.mu, mesh.rho))
    vp_e = np.sqrt((kappa + 4.0 / 3.0 * mu) / r).mean()
    vs_e = np.sqrt(mu / r).mean()
    u = np.zeros((o.nglob, 3))
    for comp, imp in ((2, rho * vp_e), (0, rho * vs_e), (1, rho * vs_e)):
        v = np.zeros((o.nglob, 3))
        v[:, comp] = 1.0
        a = o.apply_stiffness(u, v, with_stacey=True)
        assert abs(a[:, comp].sum() + imp * area) < 1e-6 * imp * area
        others = [c for c in range(3) if c != comp]
        assert np.abs(a[:, others].sum(axis=0)).max() < 1e-9 * imp * area
