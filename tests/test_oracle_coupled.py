"""Pins the coupled elastic-acoustic oracle (CPU only) on the reference's two
coupled SPECFEM2D fixtures with the reference's criterion (their config.yaml
tolerance is 1e-3): water above the solid (restated from the fixture's
Par_file: the fixture ships no database) and water below the solid (the
reference mesher's own database).  Stations lie in both media.  A sign
flipped in the solid-side coupling term, planted in the oracle on the CPU,
must fail the comparison."""

import numpy as np

from tests._cases import coupled_fixture_case, coupled_variant2_case, reference_metric


def _run(oracle_lib, cfg, scales=None):
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    assert o.coupled
    if scales:
        o.plant_coupling_scales(*scales)
    return o, o.run(types=cfg.seismo_types)


def test_water_above_solid_matches_specfem2d(oracle_lib):
    cfg, traces = coupled_fixture_case()
    o, s = _run(oracle_lib, cfg)
    assert abs(o.t0 - (-0.12)) < 1e-6
    assert reference_metric(s, traces) < 1e-4


def test_water_below_solid_from_database_matches_specfem2d(oracle_lib):
    cfg, traces = coupled_variant2_case()
    # the database's fluid-solid pairs are the edge-sharing (fluid, solid) pairs
    m = cfg.mesh
    pairs = set(map(tuple, cfg.meta["database"]["fluid_solid"]))
    corner = {}
    for e in range(m.nspec):
        c = m.knods[e, :4]
        for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
            corner.setdefault(tuple(sorted((c[a], c[b]))), []).append(e)
    shared = {(a, b) if m.acoustic[a] else (b, a) for es in corner.values() if len(es) == 2
              for a, b in [es] if m.acoustic[a] != m.acoustic[b]}
    assert pairs == shared and len(pairs) == 144
    _, s = _run(oracle_lib, cfg)
    assert reference_metric(s, traces) < 1e-4


def test_planted_coupling_sign_error_is_caught(oracle_lib):
    cfg, traces = coupled_variant2_case()
    _, s = _run(oracle_lib, cfg, scales=(1.0, -1.0))
    assert reference_metric(s, traces) > 1e-2
