"""Packs the SPECFEM2D golden traces that the reference's displacement tests
compare against into small .npz fixtures (data only).

Source directories (read as text, in this container only):
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomain
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomainStaceyDirichletBC
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousAcousticDomain
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousAcousticDomainDirichletBC
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousAcousticDomainCompositeStaceyDirichletBC
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/AcousticElasticCoupledDomain
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/AcousticElasticCoupledDomainVariant2
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomainStaceyBC
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomainWAdjacencyMap
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomainMomentTensor
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/HomogeneousElasticDomainMomentTensorHeaviside

Each case's mesh database (database.bin, written by the reference's mesher)
is also copied, gzip-compressed, as specfem2d_<case>.database.bin.gz (the
coupled case ships no database; its mesh is restated from its Par_file).

Each fixture holds the station table (STATIONS: name, network, x, z), the
sample times, and traces[type][station][component][t] for the seismogram types
present (d, v, a; pressure 'p' traces are stored separately).  The problem set
up (80x60 Q9 box, dt=1.1e-3, 300 or 500 steps, Ricker, Dirac or Heaviside
force and moment-tensor sources) is restated in tests/_cases.py from the
fixtures' Par_file / sources.yaml.

Usage: python tests/golden/make_specfem2d_traces.py [/root/reference]
"""

import gzip
import os
import sys

import numpy as np

BASE = "tests/unit-tests/displacement_tests/Newmark/serial/dim2"
CASES = ["HomogeneousElasticDomain", "HomogeneousElasticDomainStaceyDirichletBC",
         "HomogeneousAcousticDomain", "HomogeneousAcousticDomainDirichletBC",
         "HomogeneousAcousticDomainCompositeStaceyDirichletBC", "AcousticElasticCoupledDomain",
         "AcousticElasticCoupledDomainVariant2", "HomogeneousElasticDomainStaceyBC",
         "HomogeneousElasticDomainWAdjacencyMap", "HomogeneousElasticDomainMomentTensor",
         "HomogeneousElasticDomainMomentTensorHeaviside"]


def main(ref="/root/reference"):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for case in CASES:
        d = os.path.join(ref, BASE, case)
        if os.path.exists(os.path.join(d, "database.bin")):  # the coupled case ships none
            with open(os.path.join(d, "database.bin"), "rb") as fh:
                db = fh.read()
            with open(os.path.join(out_dir, f"specfem2d_{case}.database.bin.gz"), "wb") as fh:
                fh.write(gzip.compress(db, 9, mtime=0))
        names, nets, xs, zs = [], [], [], []
        with open(os.path.join(d, "STATIONS")) as fh:
            for line in fh:
                p = line.split()
                if len(p) >= 4:
                    names.append(p[0])
                    nets.append(p[1])
                    xs.append(float(p[2]))
                    zs.append(float(p[3]))
        tdir = os.path.join(d, "traces")
        files = set(os.listdir(tdir))
        types = [t for t in "dva" if f"{nets[0]}.{names[0]}.S2.BXX.sem{t}" in files]
        times = None
        traces = []
        for t in types:
            per_sta = []
            for n, net in zip(names, nets):
                comps = []
                for comp in ("BXX", "BXZ"):
                    arr = np.loadtxt(os.path.join(tdir, f"{net}.{n}.S2.{comp}.sem{t}"))
                    times = arr[:, 0] if times is None else times
                    comps.append(arr[:, 1])
                per_sta.append(comps)
            traces.append(per_sta)
        extra = {}
        if f"{nets[0]}.{names[0]}.S2.PRE.semp" in files:
            extra["pressure"] = np.array([
                np.loadtxt(os.path.join(tdir, f"{net}.{n}.S2.PRE.semp"))[:, 1]
                for n, net in zip(names, nets)])
        np.savez_compressed(os.path.join(out_dir, f"specfem2d_{case}.npz"),
                            names=np.array(names), x=np.array(xs), z=np.array(zs),
                            times=times, types=np.array(list(types)),
                            traces=np.array(traces), **extra)
        print(case, len(names), "stations, types", types, "nt", len(times))


if __name__ == "__main__":
    main(*sys.argv[1:])
