"""YAML front-end with the reference's Python entry points.

``_initialize(argv)``, ``_execute(parameter_yaml, default_yaml)`` and
``_finalize()`` mirror the reference's binding (src/python/core.cpp:10-80):
``_execute`` reads a ``specfem_config.yaml`` (as text), the sources file, the
STATIONS file and the mesh database it names, runs the forward simulation
on the GPU and writes the seismograms in the reference's ASCII format.
Keys follow the reference's parameter files (include/parameter_parser/*):

  parameters:
    header: {title, description}
    simulation-setup:
      quadrature: {quadrature-type: GLL4}
      solver: {time-marching: {time-scheme: {type: Newmark, dt, nstep}}}
      simulation-mode: {forward: {writer: {seismogram: {format: ascii,
                                                        directory}}}}
    receivers: {stations, angle, seismogram-type: [...], nstep_between_samples}

Enumerated strings compare case-insensitively with the reference's aliases
(core/specfem/macros/config_strings.hpp): ``ascii``/``txt``,
``displacement``/``disp``/``displ``/``d`` and the like.  The writer's
``directory`` defaults to ``base_dir`` and must exist, its ``format`` to
ASCII (src/parameter_parser/writer/seismogram.cpp:11-56); ``output-format``
and ``output-folder`` are read as aliases.  The receiver ``angle`` is kept
and not applied, as the reference stores it without using it
(core/specfem/assembly/receivers/dim2/receivers.cpp:97-101).
    databases: {mesh-database}
    sources: <path to a sources YAML>

Sources YAML: ``sources: [{force: {x, z, angle, Ricker: {factor, tshift,
f0}} or Dirac: {factor, tshift}}, {moment-tensor: {x, z, Mxx, Mzz, Mxz,
Heaviside: {hdur, factor, tshift}}}, ...]`` (include/parameter_parser/
sources.hpp, core/specfem/source/source.tpp:7-60).  Relative paths are
resolved against ``base_dir`` (the reference resolves them against the
working directory).  Only what this build implements is accepted (elastic
isotropic P-SV or acoustic isotropic meshes, GLL4, Newmark, forward mode,
force and moment-tensor sources with Ricker, Dirac or Heaviside time
functions, ASCII seismograms of displacement, velocity, acceleration and
pressure); anything else raises ``ValueError`` naming the key.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import yaml

from .config import (SRC_FORCE, SRC_MOMENT, STF_DIRAC, STF_HEAVISIDE, STF_RICKER, Config2D,
                     Receivers2D, Sources2D)
from .meshdb import read_mesh_database_2d
from .seismo_io import write_seismograms

# seismogram types and their aliases (config_strings.hpp:25-58), lower case
_SEISMO_ALIASES = {
    "displacement": ("displacement", "disp", "displ", "d"),
    "velocity": ("velocity", "vel", "v", "veloc"),
    "acceleration": ("acceleration", "acc", "a", "accel"),
    "pressure": ("pressure", "p", "pres"),
}
_UNBUILT_SEISMO = ("rotation", "r", "rot", "intrinsic_rotation", "intrinsic_rot",
                   "intrinsic_r", "intrinsic_rotat", "intrinsic-rotation", "ir", "curl", "crl")
_ASCII = ("ascii", "txt")
_SU = ("su", "seismic_unix", "seismic-unix")
_GLL4 = ("gll4", "gll-4", "gll_4")


def seismogram_type(name):
    """The wavefield type a seismogram-type string names
    (src/parameter_parser/receivers.cpp:14-48)."""
    low = str(name).lower()
    for t, aliases in _SEISMO_ALIASES.items():
        if low in aliases:
            return t
    if low in _UNBUILT_SEISMO:
        raise ValueError(f"seismogram-type {name!r}: rotation and curl seismograms belong to "
                         "media this build does not implement")
    raise ValueError(f"seismogram-type {name!r}: unknown seismogram type")


def writer_settings(node, base_dir="."):
    """(format, directory) of a seismogram writer node as the reference
    reads it: ``directory`` (default: the working directory, here
    ``base_dir``) must exist; ``format`` (default ASCII) must name ASCII —
    Seismic Unix is declared but not implemented there either."""
    node = node or {}
    fmt = str(node.get("format", node.get("output-format", "ASCII")))
    folder = str(node.get("directory", node.get("output-folder", ".")))
    folder = folder if os.path.isabs(folder) else os.path.join(base_dir, folder)
    if not os.path.isdir(folder):
        raise ValueError(f"Output folder : {folder} does not exist.")
    if fmt.lower() in _SU:
        raise ValueError("Seismic Unix format not implemented yet")
    if fmt.lower() not in _ASCII:
        raise ValueError(f"seismogram format {fmt!r}: unknown seismogram format")
    return "ascii", folder


@dataclass
class Stations:
    names: list
    networks: list
    x: np.ndarray
    z: np.ndarray


@dataclass
class Program:
    """A parsed run: the configuration and where its seismograms go."""
    config: Config2D
    stations: Stations
    output_folder: str
    title: str = ""
    receiver_angle: float = 0.0
    extra: dict = field(default_factory=dict)


def read_stations(path):
    """STATIONS file: name, network, x, z, (elevation, burial) per line."""
    names, nets, xs, zs = [], [], [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if not p:
                continue
            if len(p) < 4:
                raise ValueError(f"{path}: station line needs name, network, x, z: {line!r}")
            names.append(p[0])
            nets.append(p[1])
            xs.append(float(p[2]))
            zs.append(float(p[3]))
    return Stations(names, nets, np.array(xs), np.array(zs))


def _get(d, *keys, default=None, required=True):
    cur = d
    for k in keys:
        if not isinstance(cur, dict) or k not in cur:
            if required:
                raise ValueError("missing key " + "/".join(keys))
            return default
        cur = cur[k]
    return cur


def read_sources(node):
    """Force and moment-tensor sources with Ricker, Dirac or Heaviside time
    functions from a sources YAML node (src/io/sources/dim2/read_sources.cpp;
    a Dirac's f0 is 1 / (10 dt), set by the context; a Heaviside's half
    duration travels in f0)."""
    items = node.get("sources", [])
    n_decl = node.get("number-of-sources")
    if n_decl is not None and int(n_decl) != len(items):
        raise ValueError(f"number-of-sources is {n_decl} but {len(items)} are listed")
    x, z, ang, f0, ts, fac, kind, skind, mom = [], [], [], [], [], [], [], [], []
    for k, item in enumerate(items):
        if "force" in item:
            f = item["force"]
            skind.append(SRC_FORCE)
            mom.append((0.0, 0.0, 0.0))
        elif "moment-tensor" in item:
            f = item["moment-tensor"]
            skind.append(SRC_MOMENT)
            mom.append((float(f["Mxx"]), float(f["Mzz"]), float(f["Mxz"])))
        else:
            raise ValueError(f"source {k}: only force and moment-tensor sources are "
                             f"supported ({list(item)})")
        if "Ricker" in f:
            r = f["Ricker"]
            kind.append(STF_RICKER)
            f0.append(float(r["f0"]))
        elif "Dirac" in f:
            r = f["Dirac"]
            kind.append(STF_DIRAC)
            f0.append(1.0)
        elif "Heaviside" in f:
            r = f["Heaviside"]
            kind.append(STF_HEAVISIDE)
            f0.append(float(r["hdur"]))
        else:
            raise ValueError(f"source {k}: only Ricker, Dirac and Heaviside time functions "
                             "are supported")
        x.append(float(f["x"]))
        z.append(float(f["z"]))
        ang.append(float(f.get("angle", 0.0)))
        ts.append(float(r.get("tshift", 0.0)))
        fac.append(float(r["factor"]))
    return Sources2D(x=np.array(x), z=np.array(z), angle=np.array(ang, np.float32),
                     f0=np.array(f0, np.float32), tshift=np.array(ts, np.float32),
                     factor=np.array(fac, np.float32), stf=np.array(kind, np.int32),
                     kind=np.array(skind, np.int32),
                     moment=np.array(mom, np.float32).reshape(-1, 3))


def parse_config(params, base_dir="."):
    """A Program from a parsed specfem_config (dict with a 'parameters'
    root, as the reference's files have)."""
    p = params.get("parameters", params)

    def path(rel):
        return rel if os.path.isabs(rel) else os.path.join(base_dir, rel)

    setup = _get(p, "simulation-setup")
    qt = _get(setup, "quadrature", "quadrature-type", required=False, default="GLL4")
    if str(qt).lower() not in _GLL4:
        raise ValueError(f"quadrature-type {qt!r}: only GLL4 (NGLL = 5) is built")
    ts = _get(setup, "solver", "time-marching", "time-scheme")
    if str(ts.get("type", "Newmark")).lower() != "newmark":
        raise ValueError(f"time-scheme type {ts.get('type')!r}: only Newmark is built")
    dt, nstep = float(ts["dt"]), int(ts["nstep"])
    mode = _get(setup, "simulation-mode")
    if "forward" not in mode:
        raise ValueError(f"simulation-mode {list(mode)}: only forward runs are driven from YAML "
                         "(gradients: specfempp_amd.adjoint)")
    seis_w = _get(mode, "forward", "writer", "seismogram", required=False, default={}) or {}
    _, out = writer_settings(seis_w, base_dir)

    rec = _get(p, "receivers")
    st = read_stations(path(rec["stations"]))
    types = tuple(seismogram_type(t) for t in rec.get("seismogram-type", ["displacement"]))
    nsbs = int(rec.get("nstep_between_samples", 1))
    angle = float(rec.get("angle", 0.0))  # stored, not applied (receivers.cpp:97-101)

    mesh, info = read_mesh_database_2d(path(_get(p, "databases", "mesh-database")))
    src_node = p.get("sources")
    if src_node is None:  # setup.cpp:50-55
        raise ValueError("Error reading specfem source configuration: no parameters/sources")
    if isinstance(src_node, str):
        with open(path(src_node)) as fh:
            src_node = yaml.safe_load(fh)
    if not isinstance(src_node, dict):
        raise ValueError("sources: expected a path or a mapping")
    sources = read_sources(src_node)
    receivers = Receivers2D(x=st.x, z=st.z, names=tuple(st.names))
    title = str(_get(p, "header", "title", required=False, default=""))
    cfg = Config2D(title or "yaml", mesh, sources, receivers, dt, nstep, nsbs=nsbs,
                   seismo_types=types, meta={"database": info})
    return Program(cfg, st, out, title, receiver_angle=angle)


class _State:
    program = None
    initialized = False
    last = None


def _initialize(argv=None):
    """Loads the GPU library (``argv`` as the reference's; unused).  False if
    already initialized."""
    if _State.initialized:
        return False
    from . import _lib
    _lib.load()
    _State.initialized = True
    return True


def _execute(parameter_string, default_string="", base_dir="."):
    """Runs the configuration given as YAML text (keys missing from it are
    taken from ``default_string``) and writes its seismograms; returns True
    on success, False when not initialized (as the reference)."""
    if not _State.initialized:
        return False
    params = yaml.safe_load(parameter_string) or {}
    defaults = yaml.safe_load(default_string) if default_string else {}
    merged = _merge(defaults or {}, params)
    prog = parse_config(merged, base_dir=base_dir)
    seis, ctx = run_program(prog)
    _State.last = (prog, seis, ctx)
    return True


def _finalize():
    if not _State.initialized:
        return False
    _State.initialized = False
    _State.last = None
    return True


def _merge(base, over):
    out = dict(base)
    for k, v in over.items():
        out[k] = _merge(out[k], v) if isinstance(v, dict) and isinstance(out.get(k), dict) else v
    return out


def run_program(prog: Program, precision=4):
    """Forward run of a parsed Program; writes the seismograms and returns
    (seismograms (nsig, ntypes, nrec, 2), ctx)."""
    from .solver import simulate
    seis, ctx = simulate(prog.config, precision=precision)
    write_seismograms(prog.output_folder, seis,
                      list(zip(prog.stations.networks, prog.stations.names)),
                      prog.config.seismo_types, prog.config.dt, ctx.t0, prog.config.nsbs)
    return seis, ctx


def run(config_path, precision=4):
    """Runs a specfem_config.yaml file (paths relative to its folder)."""
    with open(config_path) as fh:
        params = yaml.safe_load(fh)
    prog = parse_config(params, base_dir=os.path.dirname(os.path.abspath(config_path)))
    return run_program(prog, precision)
