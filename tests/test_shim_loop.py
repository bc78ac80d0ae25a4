"""The reference's forward time loop run in C++ through the operator shim
(include/specfempp_amd_shim.hpp): tests/cpp/reference_loop.cpp builds a
context from a case file and runs time_marching<forward>::run's call
sequence (every medium tag, the degree-of-freedom and element checks)
on the GPU.  Its seismograms must equal those of the same operator sequence
driven from Python bit for bit, and match the oracle: an elastic box with
pressure and sampling every second step, an acoustic box, and a coupled
box (both media advanced by their tagged operators)."""

import os
import subprocess
import sys

import numpy as np
import pytest

from specfempp_amd import _lib
from specfempp_amd.config import homogeneous_box
from specfempp_amd.solver import simulate
from tests._cases import acoustic_box, coupled_box, rel_l2

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "reference_loop")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _write_case(path, cfg):
    """The case file reference_loop.cpp reads (layout in its header)."""
    m, s, r = cfg.mesh, cfg.sources, cfg.receivers
    types = np.array([_lib.SEIS_TYPES[t] for t in cfg.seismo_types], np.int32)
    head = np.array([m.nspec, m.ngnod, m.nnodes, len(m.abs_ispec), len(m.free_ispec),
                     len(s), len(r), len(types), cfg.nsbs, cfg.nstep], np.int32)
    parts = [head, np.array([cfg.dt], np.float64),
             np.asarray(m.coord, np.float64), np.asarray(m.knods, np.int32),
             np.asarray(m.kappa, np.float32), np.asarray(m.mu, np.float32),
             np.asarray(m.rho, np.float32), np.asarray(m.abs_ispec, np.int32),
             np.asarray(m.abs_type, np.int32), np.asarray(m.free_ispec, np.int32),
             np.asarray(m.free_type, np.int32),
             np.asarray(s.x, np.float64), np.asarray(s.z, np.float64),
             np.asarray(s.angle, np.float32), np.asarray(s.f0, np.float32),
             np.asarray(s.tshift, np.float32), np.asarray(s.factor, np.float32),
             s.kinds().astype(np.int32), s.source_kinds().astype(np.int32),
             s.moments().astype(np.float32),
             np.asarray(r.x, np.float64), np.asarray(r.z, np.float64), types]
    with open(path, "wb") as fh:
        for a in parts:
            fh.write(np.ascontiguousarray(a).tobytes())


def _cases():
    import dataclasses
    el = homogeneous_box(nx=20, nz=15, nstep=201, nrec=6)
    el = dataclasses.replace(el, nsbs=2, seismo_types=("displacement", "velocity",
                                                       "acceleration", "pressure"))
    return {"elastic": el, "acoustic": acoustic_box(nstep=200),
            "coupled": coupled_box(nstep=200)}


@pytest.mark.parametrize("name", ["elastic", "acoustic", "coupled"])
def test_reference_loop_through_the_shim(oracle_lib, tmp_path, name):
    assert os.path.exists(DRIVER), "tests/cpp/reference_loop is built by __graft_entry__.build()"
    cfg = _cases()[name]
    _say(f"reference loop through the C++ shim: {name}, {cfg.nstep} steps")
    case, out = tmp_path / "case.bin", tmp_path / "seis.out"
    _write_case(case, cfg)
    r = subprocess.run([DRIVER, str(case), str(out)], capture_output=True, text=True,
                       timeout=120)
    _say("  " + (r.stdout + r.stderr).strip())
    assert r.returncode == 0
    py, ctx = simulate(cfg, fused=False)
    cpp = np.fromfile(out, dtype=np.float32).reshape(py.shape)
    assert np.array_equal(cpp, py)
    o32 = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    o64 = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                              precision=8)
    ref32 = o32.run(nsbs=cfg.nsbs, types=cfg.seismo_types)
    ref64 = o64.run(nsbs=cfg.nsbs, types=cfg.seismo_types)
    for it, t in enumerate(cfg.seismo_types):
        e = rel_l2(cpp[:, it], ref32[:, it])
        bound = max(1e-5, 2.0 * rel_l2(ref32[:, it], ref64[:, it]))
        _say(f"  {t}: rel L2 vs FP32 oracle {e:.3e} (bound {bound:.1e})")
        assert e <= bound, t
