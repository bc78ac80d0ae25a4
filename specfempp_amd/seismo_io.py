"""Seismograms in the reference's ASCII format, so that outputs compare
directly with SPECFEM2D / SPECFEM3D tools and with the reference's traces.

One file per station, component and type, two columns per line (time,
value), each written like C++'s ``std::scientific`` at its default precision
(src/io/seismogram/writer.cpp:130-134).  Time of sample i is
t0 + i * dt * nsbs (include/io/seismogram/seismogram_iterator.hpp:38).

File names:
* 2D P-SV: ``NET.STA.S2.BX{X,Z}.sem{d,v,a}`` and, for pressure, the single
  file ``NET.STA.S2.PRE.semp`` (src/io/seismogram/writer.cpp:20-60);
* 3D: ``NET.STA.S3.{band}X{X,Y,Z}.sem{d,v,a}``, with the band code from the
  sampling interval (src/io/seismogram/impl/channel_generator.cpp:3-80).
"""

from __future__ import annotations

import os

import numpy as np

EXT = {"displacement": "semd", "velocity": "semv", "acceleration": "sema",
       "pressure": "semp"}


def band_code(dt):
    """SEED band letter of a sampling interval (channel_generator.cpp:3-22)."""
    if dt >= 1.0:
        return "L"
    if dt >= 0.1:
        return "M"
    if dt > 0.0125:
        return "B"
    if dt > 0.004:
        return "H"
    if dt > 0.001:
        return "C"
    return "F"


def station_filenames(network, station, seismo_type, dim=2, dt=None):
    """File names (without folder) of one station's components."""
    ext = EXT[seismo_type]
    if dim == 2 and seismo_type == "pressure":
        return [f"{network}.{station}.S2.PRE.{ext}"]
    if dim == 2:
        return [f"{network}.{station}.S2.BX{c}.{ext}" for c in "XZ"]
    band = band_code(dt)
    return [f"{network}.{station}.S3.{band}X{c}.{ext}" for c in "XYZ"]


def sample_times(nsig, dt, t0, nsbs=1):
    return t0 + np.arange(nsig) * (dt * nsbs)


def write_seismograms(folder, seis, stations, seismo_types, dt, t0, nsbs=1):
    """Writes seis (nsig, ntypes, nrec, ncomp), ncomp 2 (2D) or 3 (3D), for
    the stations [(network, name), ...]; returns the written paths.  A 2D
    pressure type writes component 0 only (one file per station)."""
    seis = np.asarray(seis)
    nsig, ntypes, nrec, ncomp = seis.shape
    if ncomp not in (2, 3) or nrec != len(stations) or ntypes != len(seismo_types):
        raise ValueError("seismograms, stations and types disagree")
    dim = 2 if ncomp == 2 else 3
    os.makedirs(folder, exist_ok=True)
    t = sample_times(nsig, dt, t0, nsbs)
    paths = []
    for it, st in enumerate(seismo_types):
        for r, (net, sta) in enumerate(stations):
            for c, name in enumerate(station_filenames(net, sta, st, dim, dt * nsbs)):
                path = os.path.join(folder, name)
                with open(path, "w") as fh:
                    fh.writelines(f"{a:e} {b:e}\n"
                                  for a, b in zip(t, seis[:, it, r, c].astype(np.float64)))
                paths.append(path)
    return paths


def read_trace(path):
    """(times, values) of one two-column trace file."""
    a = np.loadtxt(path, ndmin=2)
    return a[:, 0], a[:, 1]


def read_seismograms(folder, stations, seismo_types, dim=2, dt=None):
    """Reads traces written by the reference or by ``write_seismograms``:
    returns (times, seis (nsig, ntypes, nrec, ncomp)); a pressure type fills
    component 0."""
    out, times = None, None
    for it, st in enumerate(seismo_types):
        for r, (net, sta) in enumerate(stations):
            for c, name in enumerate(station_filenames(net, sta, st, dim, dt)):
                tt, v = read_trace(os.path.join(folder, name))
                if out is None:
                    times = tt
                    out = np.zeros((len(tt), len(seismo_types), len(stations), dim))
                out[:, it, r, c] = v
    return times, out


def reference_error(ours, theirs, per_type=False):
    """The reference's criterion: per station, sum|ours - theirs| over every
    seismogram type, component and sample divided by sum|theirs| over the
    same; the worst station (tests/unit-tests/displacement_tests/Newmark/
    dim2/newmark_tests.cpp:189-198,337-362).  ``per_type=True`` takes each
    (type, station) pair on its own instead, which is never looser.  Both
    arrays (nsig, ntypes, nrec, ncomp)."""
    a = np.asarray(ours, dtype=np.float64)
    b = np.asarray(theirs, dtype=np.float64)
    axes = (0, 3) if per_type else (0, 1, 3)
    num = np.abs(a - b).sum(axis=axes)
    den = np.abs(b).sum(axis=axes)
    return float((num / den).max())
