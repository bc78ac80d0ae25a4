"""Pins the oracle's moment-tensor sources, Heaviside time function and
four more elastic fixtures against the SPECFEM2D traces the reference's own
displacement tests use (reference criterion per station and type,
newmark_tests.cpp:337-362), on the reference mesher's databases."""

import numpy as np
import pytest

from specfempp_amd.config import Sources2D
from tests._cases import DB_FIXTURES, db_fixture_case, reference_metric


def _run(oracle_lib, cfg, sources=None, precision=4):
    o = oracle_lib.Oracle2D(cfg.mesh, sources or cfg.sources, cfg.receivers, cfg.dt,
                            cfg.nstep, precision=precision)
    return o, o.run(types=cfg.seismo_types)


@pytest.mark.parametrize("name", sorted(DB_FIXTURES))
def test_oracle_matches_specfem2d_fixture(oracle_lib, name):
    cfg, traces, pinned = db_fixture_case(name)
    o, s = _run(oracle_lib, cfg)
    assert s.shape == (cfg.nstep, len(cfg.seismo_types), len(cfg.receivers), 2)
    # start time: -1.2 / f0 (Ricker) or -2 hdur (Heaviside)
    t0 = -0.2 if "Heaviside" in name else -0.12
    assert abs(o.t0 - t0) < 1e-6
    for i in pinned:
        assert reference_metric(s[:, i:i + 1], traces[i:i + 1]) < 1e-3, cfg.seismo_types[i]


@pytest.mark.parametrize("plant", ["sign", "transpose"])
def test_planted_moment_tensor_mistakes_are_caught(oracle_lib, plant):
    """Plausible mistakes in the moment tensor, planted in the oracle's input
    on the CPU, fail the golden comparison: a flipped Mxz, and Mxz read as a
    diagonal term (the tensor's rows and columns confused)."""
    cfg, traces, _ = db_fixture_case("HomogeneousElasticDomainMomentTensor")
    m = (0.0, 0.0, -1.0) if plant == "sign" else (1.0, 0.0, 0.0)
    src = Sources2D.moment_tensor([2500.0], [2500.0], *m, f0=10.0, factor=1e8)
    _, s = _run(oracle_lib, cfg, src)
    assert reference_metric(s[:, :1], traces[:1]) > 1e-2


def test_heaviside_table_values(oracle_lib):
    """The Heaviside table equals factor (1 + erf(t / (0.1 / 1.628))) / 2 on
    t = -0.2 + i dt (scipy's erf in double; FP32 table), starts near zero
    and ends at the factor."""
    from scipy.special import erf
    cfg, _, _ = db_fixture_case("HomogeneousElasticDomainMomentTensorHeaviside")
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    stf = np.asarray(o.get("stf"), dtype=np.float64).reshape(cfg.nstep, -1)[:, 0]
    t = -0.2 + np.arange(cfg.nstep) * np.float32(cfg.dt)
    want = 1e8 * 0.5 * (1.0 + erf(t / (0.1 / 1.628)))
    np.testing.assert_allclose(stf, want, rtol=2e-5, atol=1e8 * 2e-6)
    assert stf[0] < 1e8 * 1e-5 and abs(stf[-1] - 1e8) < 1e8 * 1e-6
