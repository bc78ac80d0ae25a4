"""The YAML front-end on the Stacey fixture, set up as the reference's own
test is (specfem_config.yaml + sources.yaml + STATIONS + the mesher's
database.bin): the parsed run equals the restated fixture case, and the
oracle on it meets the SPECFEM2D traces; on the GPU, _initialize / _execute
/ _finalize write reference-format seismograms that meet them too."""

import gzip
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
            format: ascii
            directory: "OUTPUT_FILES"
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
    (tmp_path / "OUTPUT_FILES").mkdir()
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
                     ("format: ascii", "format: hdf5")):
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
    (tmp_path / "OUTPUT_FILES").mkdir()
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
    assert reference_error(s, _acoustic_traces(g), per_type=True) < 5e-4


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
    err = reference_error(s, _acoustic_traces(g), per_type=True)
    print("acoustic YAML run vs SPECFEM2D (reference metric, per type):", err)
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
    (tmp_path / "OUTPUT_FILES").mkdir()
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


# ---- the reference fixtures' own run files, unchanged ----------------------
# Each case's specfem_config.yaml, sources file and STATIONS as the reference
# holds them, laid out under the reference tests' working directory
# (tests/unit-tests, which the configs' relative paths start from), with the
# case's database.bin.
REF_CASES = sorted(os.listdir(os.path.join(GOLDEN, "configs")))
REF_REL = "displacement_tests/Newmark/serial/dim2"
_LETTERS = {"d": "displacement", "v": "velocity", "a": "acceleration"}


def _reference_layout(root, case):
    d = root / REF_REL / case
    d.mkdir(parents=True)
    src = os.path.join(GOLDEN, "configs", case)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), d / f)
    with gzip.open(os.path.join(GOLDEN, f"specfem2d_{case}.database.bin.gz")) as fi, \
            open(d / "database.bin", "wb") as fo:
        shutil.copyfileobj(fi, fo)
    with open(d / "specfem_config.yaml") as fh:
        return fh.read()


def _expected_types(g):
    types = tuple(_LETTERS[str(t)] for t in g["types"])
    return types + (("pressure",) if "pressure" in g.files else ())


@pytest.mark.parametrize("case", REF_CASES)
def test_reference_config_parses_unchanged(tmp_path, case):
    """Every fixture's own specfem_config.yaml (writer keys format /
    directory, receiver angle, its sources file and STATIONS) parses
    unchanged into the run its traces were made with: stations, seismogram
    types (pressure included), step count, and the output folder '.'."""
    text = _reference_layout(tmp_path, case)
    prog = program.parse_config(yaml.safe_load(text), base_dir=str(tmp_path))
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{case}.npz"))
    cfg = prog.config
    np.testing.assert_allclose(cfg.receivers.x, g["x"])
    np.testing.assert_allclose(cfg.receivers.z, g["z"])
    assert list(prog.stations.names) == [str(n) for n in g["names"]]
    assert cfg.seismo_types == _expected_types(g)
    assert cfg.nstep == len(g["times"])
    assert os.path.samefile(prog.output_folder, tmp_path)
    assert prog.receiver_angle == 0.0


def test_writer_keys_and_aliases(tmp_path):
    """The writer as the reference reads it: format compared without case
    (ASCII / txt), Seismic Unix declared but not implemented, an unknown
    format and a missing directory refused; seismogram types by their
    aliases; a nonzero receiver angle kept and not applied."""
    (tmp_path / "out").mkdir()
    assert program.writer_settings({"format": "ASCII", "directory": "out"}, str(tmp_path)) == (
        "ascii", str(tmp_path / "out"))
    assert os.path.samefile(program.writer_settings({"format": "txt"}, str(tmp_path))[1],
                            tmp_path)
    assert program.writer_settings({"output-folder": "out"}, str(tmp_path))[1].endswith("out")
    with pytest.raises(ValueError, match="Seismic Unix"):
        program.writer_settings({"format": "SU"}, str(tmp_path))
    with pytest.raises(ValueError, match="unknown seismogram format"):
        program.writer_settings({"format": "HDF5"}, str(tmp_path))
    with pytest.raises(ValueError, match="does not exist"):
        program.writer_settings({"directory": "missing"}, str(tmp_path))
    assert [program.seismogram_type(t) for t in ("Displ", "VEL", "a", "PRES")] == [
        "displacement", "velocity", "acceleration", "pressure"]
    with pytest.raises(ValueError, match="rotation and curl"):
        program.seismogram_type("rotation")
    text = _reference_layout(tmp_path, "HomogeneousElasticDomainMomentTensor")
    prog = program.parse_config(yaml.safe_load(text.replace("angle: 0.0", "angle: 30.0")),
                                base_dir=str(tmp_path))
    assert prog.receiver_angle == 30.0


def test_reference_config_run_meets_specfem2d(oracle_lib, tmp_path):
    """The moment-tensor fixture's own files, unchanged, run on the oracle:
    every type (d, v, a and the elastic pressure) meets the SPECFEM2D traces
    by the reference's pooled per-station criterion."""
    case = "HomogeneousElasticDomainMomentTensor"
    text = _reference_layout(tmp_path, case)
    prog = program.parse_config(yaml.safe_load(text), base_dir=str(tmp_path))
    cfg = prog.config
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{case}.npz"))
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    s = o.run(types=cfg.seismo_types)
    theirs = np.transpose(g["traces"], (3, 0, 1, 2))
    p = np.zeros(theirs.shape[:1] + (1,) + theirs.shape[2:])
    p[:, 0, :, 0] = g["pressure"].T
    theirs = np.concatenate([theirs, p], axis=1)
    assert reference_error(s, theirs) < 1e-3
    assert reference_error(s, theirs, per_type=True) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["HomogeneousElasticDomain",
                                  "HomogeneousElasticDomainMomentTensor",
                                  "HomogeneousElasticDomainMomentTensorHeaviside"])
def test_reference_config_execute_with_elastic_pressure(tmp_path, case):
    """The three elastic fixtures that record pressure, run through
    _initialize / _execute / _finalize from their own files unchanged: the
    written d/v/a/PRE files meet the SPECFEM2D traces by the reference's
    pooled criterion, and the pressure alone by 1e-3."""
    text = _reference_layout(tmp_path, case)
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{case}.npz"))
    assert program._initialize([]) is True
    try:
        assert program._execute(text, "", base_dir=str(tmp_path)) is True
    finally:
        assert program._finalize() is True
    types = _expected_types(g)
    st = [("AA", str(n)) for n in g["names"]]
    t, s = read_seismograms(tmp_path, st, types)
    np.testing.assert_allclose(t, g["times"], atol=1e-6)
    theirs = np.transpose(g["traces"], (3, 0, 1, 2))
    p = np.zeros(theirs.shape[:1] + (1,) + theirs.shape[2:])
    p[:, 0, :, 0] = g["pressure"].T
    theirs = np.concatenate([theirs, p], axis=1)
    pooled = reference_error(s, theirs)
    pres = reference_error(s[:, -1:], theirs[:, -1:], per_type=True)
    print(f"{case}: pooled {pooled:.3e}, pressure {pres:.3e}", flush=True)
    assert pooled < 1e-3 and pres < 1e-3
