"""Pins the acoustic half of the CPU oracle (CPU only).

* The reference's three acoustic SPECFEM2D fixtures, read from the
  reference mesher's databases, with the reference's criterion
  (newmark_tests.cpp:337-362; the fixtures' config.yaml tolerance is 1e-3):
  Dirac source with free surfaces on all sides (d, v, a and pressure), four
  Ricker sources with free surfaces on all sides, and the same with Stacey
  on three sides and a free top.
* Planted mistakes must fail those comparisons: a flipped source, a missing
  free surface, a missing Stacey condition.
* Closed forms: without Stacey the mass is sum over elements of area / kappa;
  a tagged free-surface point with no source in its elements never moves.
"""

import dataclasses

import numpy as np
import pytest

from tests._cases import ACOUSTIC_FIXTURES, acoustic_box, acoustic_fixture_case, reference_metric

BOUNDS = {"HomogeneousAcousticDomain": 5e-4,
          "HomogeneousAcousticDomainDirichletBC": 1e-4,
          "HomogeneousAcousticDomainCompositeStaceyDirichletBC": 1e-4}


def _run(oracle_lib, cfg, precision=4):
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=precision)
    return o, o.run(types=cfg.seismo_types)


@pytest.mark.parametrize("name", sorted(ACOUSTIC_FIXTURES))
def test_oracle_matches_specfem2d_acoustic(oracle_lib, name):
    cfg, traces = acoustic_fixture_case(name)
    o, s = _run(oracle_lib, cfg)
    g = np.load(f"tests/golden/specfem2d_{name}.npz")
    np.testing.assert_allclose(o.t0 + cfg.dt * np.arange(cfg.nstep), g["times"], atol=1e-6)
    for it, t in enumerate(cfg.seismo_types):
        e = reference_metric(s[:, it:it + 1], traces[it:it + 1])
        assert e < BOUNDS[name], (t, e)


def test_planted_mistakes_fail_the_acoustic_fixtures(oracle_lib):
    # flipped Dirac source
    cfg, traces = acoustic_fixture_case("HomogeneousAcousticDomain")
    src = dataclasses.replace(cfg.sources, factor=-cfg.sources.factor)
    _, s = _run(oracle_lib, dataclasses.replace(cfg, sources=src, seismo_types=("pressure",)))
    assert reference_metric(s, traces[3:4]) > 1.0
    # no free surface: the Dirichlet fixture's reflections change sign
    cfg, traces = acoustic_fixture_case("HomogeneousAcousticDomainDirichletBC")
    mesh = dataclasses.replace(cfg.mesh, free_ispec=np.zeros(0, np.int32),
                               free_type=np.zeros(0, np.int32))
    _, s = _run(oracle_lib, dataclasses.replace(cfg, mesh=mesh))
    assert reference_metric(s, traces) > 1e-2
    # no Stacey: the composite fixture's absorbing sides reflect
    cfg, traces = acoustic_fixture_case("HomogeneousAcousticDomainCompositeStaceyDirichletBC")
    mesh = dataclasses.replace(cfg.mesh, abs_ispec=np.zeros(0, np.int32),
                               abs_type=np.zeros(0, np.int32))
    _, s = _run(oracle_lib, dataclasses.replace(cfg, mesh=mesh))
    assert reference_metric(s, traces) > 1e-2


def test_acoustic_mass_closed_form(oracle_lib):
    """No Stacey: sum of the mass over the points = sum over elements of
    area / kappa (quadrature of 1/kappa, exact for element-constant kappa on
    affine elements)."""
    cfg = acoustic_box(nx=6, nz=5, absorbing=(), nstep=2)
    kappa_e = cfg.mesh.kappa[:, 2, 2].astype(np.float64)
    cfg.mesh.kappa[:] = cfg.mesh.kappa[:, 2:3, 2:3]  # element-constant
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=8)
    minv = o.get("minv").reshape(-1, 2)[:, 0]
    area = 50.0 * 50.0
    np.testing.assert_allclose((1.0 / minv).sum(), (area / kappa_e).sum(), rtol=1e-12)


def test_free_surface_points_stay_at_rest(oracle_lib):
    """Tagged free-surface points get no element force; with no source
    touching them their potential stays exactly zero."""
    cfg = acoustic_box(nx=10, nz=8, nstep=200, free=("top", "left"))
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=8)
    o.run(types=("displacement",))
    u = o.get("u").reshape(-1, 2)[:, 0]
    ibool = o.get("ibool").astype(np.int64).reshape(cfg.mesh.nspec, 5, 5)
    top = ibool[cfg.mesh.free_ispec[cfg.mesh.free_type == 3], 4, :].ravel()
    left = ibool[cfg.mesh.free_ispec[cfg.mesh.free_type == 4], :, 0].ravel()
    assert np.abs(u).max() > 0
    assert np.all(u[top] == 0) and np.all(u[left] == 0)


def test_dirac_time_function(oracle_lib):
    """Dirac: factor * (-exp(-a t^2) / (2a)), a = pi^2 f0^2, f0 = 1/(10 dt),
    t0 = -1.2 / f0 (dirac.cpp:23-52, impl/time_functions.hpp:9-14)."""
    cfg = acoustic_box(nx=6, nz=5, nstep=50)
    cfg.sources.stf[:] = 1
    cfg.sources.tshift[:] = 0.0
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=8)
    dt = float(np.float32(cfg.dt))  # the oracle takes dt as FP32
    f0 = 1.0 / (10.0 * dt)
    assert abs(o.t0 - (-1.2 / f0)) < 1e-12
    t = o.t0 + dt * np.arange(cfg.nstep)
    a = np.pi ** 2 * f0 ** 2
    want = 1e10 * (-np.exp(-a * t * t) / (2 * a))
    stf = o.get("stf").reshape(cfg.nstep, len(cfg.sources), 2)[:, 0, 0]
    np.testing.assert_allclose(stf, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
