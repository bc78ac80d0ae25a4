"""CPU-side checks of the C ABI library and the host logic (no GPU calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

from specfempp_amd import _lib
from specfempp_amd.config import homogeneous_box, layered_box
from specfempp_amd.mesh import NGLL, box_mesh, elastic_moduli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "specfempp_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smx_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    # host-only calls that touch no device
    lib.smx_version.restype = ctypes.c_char_p
    assert lib.smx_version().startswith(b"specfempp_amd")
    lib.smx_ctx_destroy.argtypes = [ctypes.c_void_p]
    assert lib.smx_ctx_destroy(None) == 0
    lib.smx_initialize.argtypes = [ctypes.c_void_p]
    assert lib.smx_initialize(None) == -1  # null context is an argument error
    lib.smx_last_error.restype = ctypes.c_char_p
    assert b"null context" in lib.smx_last_error()


def test_cpp_shim_compiles_and_iterates(tmp_path):
    """The reference-named C++ classes compile against the C header, and the
    time scheme's iteration and seismogram cadence (timescheme.hpp:16-201)
    run as the reference's do: forward steps 0..nstep-1, backward
    nstep-1..0, a sample every nstep_between_samples steps."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include",
                        f"{ROOT}/tests/cpp/reference_loop.cpp"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    src = tmp_path / "iterate.cpp"
    src.write_text(
        '#include <cstdio>\n#include "specfempp_amd_shim.hpp"\n'
        "using namespace specfempp_amd;\n"
        "int main() { newmark n(nullptr, 7, 3, 0.5f);\n"
        "  for (const auto s : n.iterate_forward()) std::printf(\"f%d:%g \", s.istep, s.dt);\n"
        "  for (const auto s : n.iterate_backward()) std::printf(\"b%d \", s.istep);\n"
        "  for (int i = 0; i < 7; ++i) if (n.compute_seismogram(i)) n.increment_seismogram_step();\n"
        "  std::printf(\"n%d m%d p%d\", n.get_seismogram_step(), n.get_max_seismogram_step(),\n"
        "              n.apply_predictor_phase_backward(medium_tag::acoustic));\n"
        "  return 0; }\n")
    exe = tmp_path / "iterate"
    r = subprocess.run(["g++", "-std=c++17", f"-I{ROOT}/include", str(src),
                        f"-L{ROOT}/specfempp_amd", "-lspecfempp_amd", "-L/opt/rocm/lib",
                        "-Wl,-rpath-link,/opt/rocm/lib", f"-Wl,-rpath,{ROOT}/specfempp_amd",
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout
    assert out == ("f0:0.5 f1:0.5 f2:0.5 f3:0.5 f4:0.5 f5:0.5 f6:0.5 "
                   "b6 b5 b4 b3 b2 b1 b0 n3 m2 p0")


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_box_mesh_topology():
    for ngnod in (4, 9):
        m = box_mesh(0, 400, 0, 300, 8, 6, ngnod=ngnod, absorbing=("bottom", "left", "right"))
        assert m.nspec == 48 and m.ngnod == ngnod
        assert m.expected_nglob() == 33 * 25
        # absorbing edges: 8 bottom + 6 left + 6 right
        assert len(m.abs_ispec) == 20
        assert set(m.abs_type.tolist()) == {1, 2, 4}
        # corner nodes of element 0 are the box corner square
        c = m.coord[m.knods[0, :4]]
        np.testing.assert_allclose(c, [[0, 0], [50, 0], [50, 50], [0, 50]])


def test_elastic_moduli_reference_promotions():
    k, mu, rho = elastic_moduli(np.float32(2700.0), np.float32(3000.0), np.float32(1732.0))
    assert mu == np.float32(np.float32(2700.0 * 1732.0) * np.float32(1732.0))
    exact = 2700.0 * (np.float64(np.float32(9e6)) - (4.0 / 3.0) * 1732.0 * 1732.0)
    assert k == np.float32(exact)


def test_layered_config_is_seeded():
    a = layered_box(nx=40, nz=30, nstep=10)
    b = layered_box(nx=40, nz=30, nstep=10)
    np.testing.assert_array_equal(a.mesh.kappa, b.mesh.kappa)
    np.testing.assert_array_equal(a.sources.x, b.sources.x)
    assert len(set(np.unique(a.mesh.mu).tolist())) == 6
    rho = a.mesh.rho[:, 0, 0]
    kappa, mu = a.mesh.kappa[:, 0, 0].astype(np.float64), a.mesh.mu[:, 0, 0].astype(np.float64)
    vp = np.sqrt((kappa + 4 / 3 * mu) / rho)
    np.testing.assert_allclose(rho, 310 * vp ** 0.25, rtol=1e-4)
    assert vp.min() >= 2000 - 1 and vp.max() <= 6000 + 1


def test_homogeneous_config():
    c = homogeneous_box(nstep=5)
    assert c.mesh.nspec == 4800 and c.mesh.expected_nglob() == 77361
    assert len(c.receivers) == 50 and len(c.sources) == 1
    assert c.mesh.kappa.shape == (4800, NGLL, NGLL)


def _acoustic_cfg(**kw):
    from tests._cases import acoustic_box
    return acoustic_box(nx=6, nz=5, nstep=10, **kw)


def test_acoustic_mesh_model():
    from specfempp_amd.mesh import acoustic_moduli
    c = _acoustic_cfg()
    assert c.mesh.acoustic.all() and (c.mesh.mu == 0).all()
    # kappa = density * cp * cp in float (acoustic/isotropic/material.hpp:49-53)
    k, mu, rho = acoustic_moduli(np.float32(2700.0), np.float32(3000.0))
    assert k == np.float32(np.float32(2700.0 * 3000.0) * np.float32(3000.0)) and mu == 0
    # free surface: the top side's elements, type 3
    assert len(c.mesh.free_ispec) == 6 and set(c.mesh.free_type.tolist()) == {3}


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_acoustic_set_up_errors_are_host_side(monkeypatch):
    """Invalid acoustic set-ups are rejected by the host assembly, before any
    device call, with the reference's wording of what is missing."""
    import dataclasses
    from specfempp_amd.config import Receivers2D, Sources2D
    from specfempp_amd.solver import Context
    c = _acoustic_cfg()
    # a coupled mesh (one solid element at the bottom-left corner): a
    # free-surface entity on the solid element, or a moment-tensor source
    # given without its components, are rejected before any device call
    mu = c.mesh.mu.copy()
    mu[0] = 1e9
    mixed = dataclasses.replace(c.mesh, mu=mu, free_ispec=np.array([0], np.int32),
                                free_type=np.array([1], np.int32))
    with pytest.raises(_lib.SmxError, match="elastic element"):
        Context(mixed, c.sources, c.receivers, c.dt, c.nstep)
    mixed = dataclasses.replace(mixed, free_ispec=np.zeros(0, np.int32),
                                free_type=np.zeros(0, np.int32))
    import specfempp_amd.solver as solver_mod
    with monkeypatch.context() as mp:  # NULL moment array across the C ABI
        mp.setattr(solver_mod, "_moments", lambda s: np.zeros((0, 3), np.float32))
        mt = dataclasses.replace(c.sources, kind=np.array([1, 0, 0], np.int32))
        with pytest.raises(_lib.SmxError, match="moment"):
            Context(mixed, mt, c.receivers, c.dt, c.nstep)
    e = homogeneous_box(nstep=5)
    # a free surface on an elastic mesh
    fe = dataclasses.replace(e.mesh, free_ispec=np.array([0], np.int32),
                             free_type=np.array([1], np.int32))
    with pytest.raises(_lib.SmxError, match="free-surface"):
        Context(fe, e.sources, e.receivers, e.dt, e.nstep)
    # unknown time function
    bad = dataclasses.replace(c.sources, stf=np.array([0, 7, 1], np.int32))
    with pytest.raises(_lib.SmxError, match="time function"):
        Context(c.mesh, bad, c.receivers, c.dt, c.nstep)
    # batched shots are elastic only
    with pytest.raises(_lib.SmxError, match="batched"):
        Context.batch([c, c])
