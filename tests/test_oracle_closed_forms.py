"""Pins the oracle's building blocks against closed forms and invariants
that a plausible mistake (dropped term, wrong sign or index, transposed
operand) would break."""

import numpy as np
import pytest

from specfempp_amd.config import Receivers2D, Sources2D
from specfempp_amd.mesh import box_mesh


def _make(oracle_lib, nx=6, nz=4, h=50.0, absorbing=(), ngnod=4, dt=1e-3, nstep=10,
          src=(150.0, 100.0), rec=((100.0, 100.0),), precision=4, rho=2700.0, vp=3000.0,
          vs=1732.0, angle=0.0):
    mesh = box_mesh(0.0, nx * h, 0.0, nz * h, nx, nz, ngnod=ngnod, absorbing=absorbing,
                    rho=rho, vp=vp, vs=vs)
    s = Sources2D.ricker([src[0]], [src[1]], f0=10.0, factor=1e10, angle=angle)
    r = Receivers2D(x=np.array([p[0] for p in rec]), z=np.array([p[1] for p in rec]))
    return oracle_lib.Oracle2D(mesh, s, r, dt, nstep, precision=precision), mesh


def test_gll_closed_forms(oracle_lib):
    o, _ = _make(oracle_lib)
    xi, w, hp = o.get("xi"), o.get("w"), o.get("hprime").reshape(5, 5)
    s = np.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(xi, [-1, -s, 0, s, 1], atol=1e-7)
    np.testing.assert_allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], rtol=1e-7)
    assert abs(w.sum() - 2.0) < 1e-6
    # rows of h' annihilate constants; h' differentiates degree <= 4 exactly
    np.testing.assert_allclose(hp.sum(axis=1), 0.0, atol=2e-6)
    for k in range(1, 5):
        np.testing.assert_allclose(hp @ xi ** k, k * xi ** (k - 1), atol=5e-6)
    # the transpose would not: (h'^T) x != 1
    assert np.abs(hp.T @ xi - 1.0).max() > 0.1


@pytest.mark.parametrize("ngnod", [4, 9])
def test_affine_jacobian(oracle_lib, ngnod):
    o, mesh = _make(oracle_lib, ngnod=ngnod)
    # FP32 sums of shape-function derivatives: a few ulp
    np.testing.assert_allclose(o.get("jacobian"), 25.0 * 25.0, rtol=1e-5)
    np.testing.assert_allclose(o.get("xix"), 1 / 25.0, rtol=1e-5)
    np.testing.assert_allclose(o.get("gammaz"), 1 / 25.0, rtol=1e-5)
    np.testing.assert_allclose(o.get("xiz"), 0.0, atol=1e-7)
    np.testing.assert_allclose(o.get("gammax"), 0.0, atol=1e-7)
    assert o.nglob == (4 * 6 + 1) * (4 * 4 + 1)


def test_mass_sums_to_rho_times_area(oracle_lib):
    o, _ = _make(oracle_lib, rho=2000.0)
    M = 1.0 / o.get("minv").reshape(-1, 2)
    np.testing.assert_allclose(M.sum(axis=0), 2000.0 * 300.0 * 200.0, rtol=1e-5)


def test_stacey_mass_term_closed_form(oracle_lib):
    """Bottom edge only: n = (0, -L/2 per unit xi); the mass gains
    (dt/2) rho vs L in x and (dt/2) rho vp L in z."""
    rho, vp, vs, dt = 2000.0, 3000.0, 1500.0, 1e-3
    o, _ = _make(oracle_lib, absorbing=("bottom",), rho=rho, vp=vp, vs=vs, dt=dt)
    M = 1.0 / o.get("minv").reshape(-1, 2)
    L = 300.0
    extra = M.sum(axis=0) - rho * 300.0 * 200.0
    np.testing.assert_allclose(extra, [0.5 * dt * rho * vs * L, 0.5 * dt * rho * vp * L],
                               rtol=2e-4)
    n1 = o.get("stacey_n1").reshape(-1, 5, 5)
    flag = o.get("stacey_flag").reshape(-1, 5, 5)
    assert flag[:, 0, :].sum() == 6 * 5 and flag[:, 1:, :].sum() == 0
    np.testing.assert_allclose(n1[:6, 0, :], -25.0, rtol=1e-6)


def test_stiffness_rigid_motion_and_symmetry(oracle_lib):
    o, mesh = _make(oracle_lib, nx=5, nz=4, precision=8)
    n = o.nglob
    # rigid translation produces no force
    o.set_field("u", np.tile([1.0, -2.0], n))
    a = o.apply_stiffness(with_stacey=False)
    assert np.abs(a).max() < 1e-12 * 2700.0 * 3000.0 ** 2  # FP64 rounding of K*u
    # K is symmetric: r2 . K r1 == r1 . K r2 (a transposed h' breaks this)
    rng = np.random.default_rng(0)
    r1, r2 = rng.standard_normal(2 * n), rng.standard_normal(2 * n)
    o.set_field("u", r1)
    a1 = o.apply_stiffness(with_stacey=False).copy()
    o.set_field("u", r2)
    a2 = o.apply_stiffness(with_stacey=False).copy()
    assert abs(r2 @ a1 - r1 @ a2) < 1e-10 * abs(r1 @ a1)
    # and -K is negative definite away from rigid motions
    assert r1 @ a1 < 0 and r2 @ a2 < 0


def test_uniform_strain_interior_forces_vanish(oracle_lib):
    """u_x = eps x gives uniform stress: interior nodal forces are zero, and
    the right/left edges carry +-sigma_xx * L_edge in total."""
    rho, vp, vs = 2000.0, 3000.0, 1500.0
    o, mesh = _make(oracle_lib, nx=6, nz=4, rho=rho, vp=vp, vs=vs, precision=8)
    from specfempp_amd.mesh import gll_coordinates
    xy = gll_coordinates(mesh)
    ib = o.get("ibool").astype(int).reshape(-1, 5, 5)
    n = o.nglob
    X = np.zeros(n)
    X[ib.ravel()] = xy[..., 0].ravel()
    Z = np.zeros(n)
    Z[ib.ravel()] = xy[..., 1].ravel()
    eps = 1e-3
    u = np.zeros((n, 2))
    u[:, 0] = eps * X
    o.set_field("u", u.ravel())
    a = o.apply_stiffness(with_stacey=False).reshape(-1, 2)
    interior = (X > 1e-6) & (X < 300 - 1e-6) & (Z > 1e-6) & (Z < 200 - 1e-6)
    lam2mu = rho * vp * vp
    lam = lam2mu - 2 * rho * vs * vs
    scale = lam2mu * eps * 200.0
    assert np.abs(a[interior]).max() < 1e-9 * scale
    right = np.abs(X - 300.0) < 1e-6
    np.testing.assert_allclose(a[right, 0].sum(), -lam2mu * eps * 200.0, rtol=1e-6)
    top = np.abs(Z - 200.0) < 1e-6
    # sigma_zz = lam*eps acts on top/bottom: net z force on the top edge
    np.testing.assert_allclose(a[top & ~right & (X > 1e-6), 1].sum(),
                               -lam * eps * 300.0 * (1 - 2 * 0.1 * 25 / 300.0),
                               rtol=1e-6)


def test_ricker_and_source_array(oracle_lib):
    dt, nstep = 1e-3, 300
    o, _ = _make(oracle_lib, dt=dt, nstep=nstep, src=(137.0, 61.0), angle=30.0)
    stf = o.get("stf").reshape(nstep, 1, 2)[:, 0, 0]
    f0, t0 = 10.0, -1.2 / 10.0
    t = t0 + dt * np.arange(nstep)
    a = np.pi ** 2 * f0 ** 2
    np.testing.assert_allclose(stf, 1e10 * (1 - 2 * a * t * t) * np.exp(-a * t * t),
                               rtol=2e-5, atol=1e10 * 2e-6)
    sa = o.get("source_array").reshape(2, 25)
    # sum of the interpolants is 1: the array sums to the force vector
    ang = np.deg2rad(30.0)
    np.testing.assert_allclose(sa.sum(axis=1), [np.sin(ang), np.cos(ang)], rtol=1e-5)


def test_receiver_interpolation_of_affine_field(oracle_lib):
    """Interpolating the displacement field u = (x, z) at a receiver returns its
    coordinates (Lagrange interpolation is exact on affine fields)."""
    pts = ((123.0, 77.0), (250.0, 150.0), (0.0, 0.0), (300.0, 200.0))
    o, mesh = _make(oracle_lib, rec=pts, precision=8)
    from specfempp_amd.mesh import gll_coordinates
    xy = gll_coordinates(mesh)
    ib = o.get("ibool").astype(int).reshape(-1, 5, 5)
    u = np.zeros((o.nglob, 2))
    u[ib.ravel()] = xy.reshape(-1, 2)
    h = o.get("receiver_h").reshape(len(pts), 5, 5)
    e = o.get("receiver_ispec").astype(int)
    got = np.einsum("rzx,rzxc->rc", h, u[ib[e]])
    np.testing.assert_allclose(got, np.array(pts), atol=1e-9)
