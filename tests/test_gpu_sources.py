"""GPU parity of four more elastic fixtures of the reference's test list on
the reference mesher's databases: Stacey on all sides, a database with the
adjacency map, and a moment-tensor source with a Ricker or a Heaviside time
function.  Each is checked against the SPECFEM2D traces with the
reference's criterion on its pinned types and against the oracle on the same
mesh."""

import sys

import numpy as np
import pytest

from specfempp_amd.solver import simulate
from tests._cases import DB_FIXTURES, db_fixture_case, pooled_metric, reference_metric, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(DB_FIXTURES))
def test_database_fixture_on_gpu(oracle_lib, name):
    cfg, traces, pinned = db_fixture_case(name)
    _say(f"database fixture {name}: GPU vs oracle and SPECFEM2D")
    ours, _ = simulate(cfg)
    ref = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt,
                              cfg.nstep).run(types=cfg.seismo_types)
    ref64 = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt,
                                cfg.nstep, precision=8).run(types=cfg.seismo_types)
    # displacement against the FP32 oracle: twice the FP32 rounding floor of
    # this case (FP32 oracle against FP64 oracle), at least 1e-5
    floor = rel_l2(ref[:, 0], ref64[:, 0])
    tol = max(1e-5, 2.0 * floor)
    e_or = rel_l2(ours[:, 0], ref[:, 0])
    e_sp = [reference_metric(ours[:, i:i + 1], traces[i:i + 1]) for i in pinned]
    pooled = pooled_metric(ours, traces)
    _say(f"  vs oracle {e_or:.3e} (bound {tol:.3e}); vs SPECFEM2D per pinned type "
         f"{np.round(e_sp, 7)}, pooled over {cfg.seismo_types} {pooled:.3e}")
    assert e_or <= tol
    assert max(e_sp) < 1e-3
    assert pooled < 1e-3
    if "pressure" in cfg.seismo_types:
        ip = cfg.seismo_types.index("pressure")
        assert rel_l2(ours[:, ip], ref[:, ip]) <= max(
            1e-5, 2.0 * rel_l2(ref[:, ip], ref64[:, ip]))
