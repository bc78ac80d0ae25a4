"""The YAML front-end on the Stacey fixture, set up as the reference's own
test is (specfem_config.yaml + sources.yaml + STATIONS + the mesher's
database.bin): the parsed run equals the restated fixture case, and the
oracle on it meets the SPECFEM2D traces; on the GPU, _initialize / _execute
/ _finalize write reference-format seismograms that meet them too."""

import os
import shutil

import numpy as np
import pytest
import yaml

from specfempp_amd import program
from specfempp_amd.seismo_io import read_seismograms, reference_error
from tests._cases import GOLDEN, fixture_case

NAME = "HomogeneousElasticDomainStaceyDirichletBC"

CONFIG = """
parameters:
  header:
    title: Homogeneous elastic domain with stacey BCs
  simulation-setup:
    quadrature:
      quadrature-type: GLL4
    solver:
      time-marching:
        time-scheme:
          type: Newmark
          dt: 1.1e-3
          nstep: 300
    simulation-mode:
      forward:
        writer:
          seismogram:
            output-format: ascii
            output-folder: "OUTPUT_FILES"
  receivers:
    stations: "fixture/STATIONS"
    angle: 0.0
    seismogram-type:
      - displacement
    nstep_between_samples: 1
  run-setup:
    number-of-processors: 1
    number-of-runs: 1
  databases:
    mesh-database: "fixture/database.bin.gz"
  sources: "fixture/sources.yaml"
"""


def _force(x, z):
    return {"force": {"x": x, "z": z, "source_surf": False, "angle": 0.0, "vx": 0.0,
                      "vz": 0.0, "Ricker": {"factor": 1e10, "tshift": 0.0, "f0": 10.0}}}


@pytest.fixture()
def run_dir(tmp_path):
    """A run folder laid out like the reference's test: STATIONS,
    sources.yaml and the database next to the config."""
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{NAME}.npz"))
    d = tmp_path / "fixture"
    d.mkdir()
    with open(d / "STATIONS", "w") as fh:
        for n, x, z in zip(g["names"], g["x"], g["z"]):
            fh.write(f"{n}    AA    {x:.7f}    {z:.7f}    0.0    0.0\n")
    src = {"number-of-sources": 4,
           "sources": [_force(500.0, 500.0), _force(3500.0, 500.0), _force(500.0, 2500.0),
                       _force(3500.0, 2500.0)]}
    with open(d / "sources.yaml", "w") as fh:
        yaml.safe_dump(src, fh)
    shutil.copy(os.path.join(GOLDEN, f"specfem2d_{NAME}.database.bin.gz"), d / "database.bin.gz")
    return tmp_path, g


def test_parsed_run_is_the_fixture_case(oracle_lib, run_dir):
    base, g = run_dir
    prog = program.parse_config(yaml.safe_load(CONFIG), base_dir=str(base))
    cfg, ref = prog.config, fixture_case(NAME)[0]
    assert (cfg.dt, cfg.nstep, cfg.nsbs, cfg.seismo_types) == (1.1e-3, 300, 1, ("displacement",))
    for a in ("x", "z", "f0", "factor", "angle", "tshift"):
        np.testing.assert_array_equal(getattr(cfg.sources, a), getattr(ref.sources, a))
    np.testing.assert_allclose(cfg.receivers.x, ref.receivers.x)
    assert prog.stations.networks[0] == "AA" and prog.output_folder.endswith("OUTPUT_FILES")
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    s = o.run(types=cfg.seismo_types)
    assert reference_error(s, np.transpose(g["traces"], (3, 0, 1, 2))) < 1e-3


def test_unsupported_settings_are_named(run_dir):
    base, _ = run_dir
    for bad, key in (("quadrature-type: GLL4", "quadrature-type: GLL7"),
                     ("type: Newmark", "type: LeapFrog"),
                     ("output-format: ascii", "output-format: hdf5")):
        with pytest.raises(ValueError, match=key.split(":")[0]):
            program.parse_config(yaml.safe_load(CONFIG.replace(bad, key)), base_dir=str(base))


def test_entry_points_need_initialize():
    program._finalize()
    assert program._execute(CONFIG) is False
    assert program._finalize() is False


@pytest.mark.gpu
def test_execute_writes_reference_seismograms(run_dir):
    base, g = run_dir
    assert program._initialize([]) is True
    try:
        assert program._initialize([]) is False
        assert program._execute(CONFIG, "", base_dir=str(base)) is True
    finally:
        assert program._finalize() is True
    st = [("AA", n) for n in g["names"]]
    t, s = read_seismograms(base / "OUTPUT_FILES", st, ("displacement",))
    np.testing.assert_allclose(t, g["times"], atol=1e-6)
    err = reference_error(s, np.transpose(g["traces"], (3, 0, 1, 2)))
    print("YAML run vs SPECFEM2D (reference metric):", err)
    assert err < 1e-3


ACOUSTIC = "HomogeneousAcousticDomain"
ACOUSTIC_TYPES = ("displacement", "velocity", "acceleration", "pressure")


@pytest.fixture()
def acoustic_run_dir(tmp_path):
    """The Dirac acoustic fixture laid out as the reference's run: one force
    source with a Dirac time function, all four seismogram types."""
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{ACOUSTIC}.npz"))
    d = tmp_path / "fixture"
    d.mkdir()
    with open(d / "STATIONS", "w") as fh:
        for n, x, z in zip(g["names"], g["x"], g["z"]):
            fh.write(f"{n}    AA    {x:.7f}    {z:.7f}    0.0    0.0\n")
    src = {"number-of-sources": 1,
           "sources": [{"force": {"x": 2500.0, "z": 2500.0, "source_surf": False,
                                  "angle": 0.0, "vx": 0.0, "vz": 0.0,
                                  "Dirac": {"factor": 1e10, "tshift": 0.0}}}]}
    with open(d / "sources.yaml", "w") as fh:
        yaml.safe_dump(src, fh)
    shutil.copy(os.path.join(GOLDEN, f"specfem2d_{ACOUSTIC}.database.bin.gz"),
                d / "database.bin.gz")
    types = "\n".join(f"      - {t}" for t in ACOUSTIC_TYPES)
    cfg = CONFIG.replace("      - displacement", types)
    return tmp_path, g, cfg


def _acoustic_traces(g):
    tr = np.transpose(g["traces"], (3, 0, 1, 2))  # (nt, 3 types, nsta, 2)
    p = np.zeros(tr.shape[:1] + (1,) + tr.shape[2:])
    p[:, 0, :, 0] = g["pressure"].T
    return np.concatenate([tr, p], axis=1)


def test_acoustic_yaml_run_is_the_fixture_case(oracle_lib, acoustic_run_dir):
    base, g, text = acoustic_run_dir
    prog = program.parse_config(yaml.safe_load(text), base_dir=str(base))
    cfg = prog.config
    assert cfg.seismo_types == ACOUSTIC_TYPES and cfg.mesh.acoustic.all()
    assert cfg.sources.kinds().tolist() == [1]
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    np.testing.assert_allclose(o.t0 + cfg.dt * np.arange(cfg.nstep), g["times"], atol=1e-6)
    s = o.run(types=cfg.seismo_types)
    assert reference_error(s, _acoustic_traces(g)) < 5e-4


@pytest.mark.gpu
def test_acoustic_execute_writes_pressure_seismograms(acoustic_run_dir):
    base, g, text = acoustic_run_dir
    assert program._initialize([]) is True
    try:
        assert program._execute(text, "", base_dir=str(base)) is True
    finally:
        assert program._finalize() is True
    st = [("AA", n) for n in g["names"]]
    assert (base / "OUTPUT_FILES" / f"AA.{g['names'][0]}.S2.PRE.semp").exists()
    t, s = read_seismograms(base / "OUTPUT_FILES", st, ACOUSTIC_TYPES)
    np.testing.assert_allclose(t, g["times"], atol=1e-6)
    err = reference_error(s, _acoustic_traces(g))
    print("acoustic YAML run vs SPECFEM2D (reference metric):", err)
    assert err < 5e-4


COUPLED = "AcousticElasticCoupledDomainVariant2"


@pytest.mark.gpu
def test_coupled_yaml_run_from_the_reference_database(tmp_path):
    """The coupled fixture laid out as the reference's run (its mesher's
    database, sources.yaml, STATIONS): _execute writes seismograms that meet
    the SPECFEM2D traces."""
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{COUPLED}.npz"))
    d = tmp_path / "fixture"
    d.mkdir()
    with open(d / "STATIONS", "w") as fh:
        for n, x, z in zip(g["names"], g["x"], g["z"]):
            fh.write(f"{n}    AA    {x:.7f}    {z:.7f}    0.0    0.0\n")
    src = {"number-of-sources": 1,
           "sources": [{"force": {"x": 1575.0, "z": 2900.0, "source_surf": False,
                                  "angle": 0.0, "vx": 0.0, "vz": 0.0,
                                  "Ricker": {"factor": 1e9, "tshift": 0.0, "f0": 10.0}}}]}
    with open(d / "sources.yaml", "w") as fh:
        yaml.safe_dump(src, fh)
    shutil.copy(os.path.join(GOLDEN, f"specfem2d_{COUPLED}.database.bin.gz"),
                d / "database.bin.gz")
    text = CONFIG.replace("dt: 1.1e-3", "dt: 0.85e-3").replace("nstep: 300", "nstep: 600")
    assert program._initialize([]) is True
    try:
        assert program._execute(text, "", base_dir=str(tmp_path)) is True
    finally:
        assert program._finalize() is True
    st = [("AA", n) for n in g["names"]]
    t, s = read_seismograms(tmp_path / "OUTPUT_FILES", st, ("displacement",))
    np.testing.assert_allclose(t, g["times"], atol=1e-6)
    err = reference_error(s, np.transpose(g["traces"], (3, 0, 1, 2)))
    print("coupled YAML run vs SPECFEM2D (reference metric):", err)
    assert err < 1e-4


def test_moment_tensor_and_heaviside_sources_are_read():
    """A reference sources.yaml with a moment tensor and a Heaviside time
    function (the HomogeneousElasticDomainMomentTensorHeaviside fixture's
    layout) reads into the moment-tensor source of that fixture."""
    import yaml
    from specfempp_amd.config import SRC_FORCE, SRC_MOMENT, STF_HEAVISIDE, STF_RICKER
    from specfempp_amd.program import read_sources
    node = yaml.safe_load("""
number-of-sources: 2
sources:
  - moment-tensor:
      x: 2500.0
      z: 2500.0
      source_surf: false
      Mxx: 0.5
      Mzz: -0.25
      Mxz: 1.0
      Heaviside:
        hdur: 0.1
        factor: 1e8
        tshift: 0.0
  - force:
      x: 100.0
      z: 200.0
      angle: 30.0
      Ricker:
        f0: 10.0
        factor: 1e10
""")
    s = read_sources(node)
    assert list(s.source_kinds()) == [SRC_MOMENT, SRC_FORCE]
    assert list(s.kinds()) == [STF_HEAVISIDE, STF_RICKER]
    np.testing.assert_allclose(s.moments()[0], [0.5, -0.25, 1.0])
    np.testing.assert_allclose(s.f0, [0.1, 10.0], rtol=1e-6)
    np.testing.assert_allclose(s.factor, [1e8, 1e10], rtol=1e-6)
    with pytest.raises(ValueError, match="time functions"):
        read_sources({"sources": [{"moment-tensor": {"x": 0, "z": 0, "Mxx": 1, "Mzz": 1,
                                                     "Mxz": 0, "External": {}}}]})
