"""Pins the CPU oracle against the SPECFEM2D golden traces the reference's
own displacement tests use, with the reference's criterion (per-station
sum|s - ref| / sum|ref| <= 1e-3, newmark_tests.cpp:351-362)."""

import numpy as np
import pytest

from tests._cases import FIXTURES, fixture_case, reference_metric


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_matches_specfem2d(oracle_lib, name):
    cfg, g = fixture_case(name)
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    assert o.nglob == (4 * 80 + 1) * (4 * 60 + 1)
    # sample times t0 + isig*dt match the fixture's time column
    np.testing.assert_allclose(o.t0 + cfg.dt * np.arange(cfg.nstep), g["times"], atol=1e-6)
    s = o.run(types=cfg.seismo_types)
    assert s.shape == (300, len(cfg.seismo_types), len(cfg.receivers), 2)
    assert reference_metric(s, g["traces"]) < 1e-3


def test_planted_sign_error_is_caught(oracle_lib):
    """A plausible mistake (one source with the wrong polarity) planted in
    the oracle's input on the CPU must fail the golden comparison: the
    criterion has teeth."""
    cfg, g = fixture_case("HomogeneousElasticDomainStaceyDirichletBC")
    src = cfg.sources
    src.factor = src.factor.copy()
    src.factor[0] = -src.factor[0]
    o = oracle_lib.Oracle2D(cfg.mesh, src, cfg.receivers, cfg.dt, cfg.nstep)
    s = o.run(types=cfg.seismo_types)
    assert reference_metric(s, g["traces"]) > 1e-2
