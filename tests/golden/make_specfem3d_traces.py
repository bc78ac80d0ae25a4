"""Packs the SPECFEM3D golden traces that the reference's 3D displacement test
compares against into a small .npz fixture (data only).

Source directory (read as text, in this container only):
  tests/unit-tests/displacement_tests/Newmark/serial/dim3/
      HomogeneousHalfspaceSmallNoABCForceSource

The fixture holds the station table (STATIONS: station, network, y, x,
elevation, z -- the column order the reference's 3D reader uses,
src/io/receivers/dim3/read_receivers.cpp), the sample times and
traces[type][station][component][t] for d, v, a and components X, Y, Z.  The
problem set-up (8x6x6 hex box of 8000x7000x7500 m, rho 2700, vp 3000,
vs 1732.051, free surfaces, force fz = 1e14 at (4000, 3500, -3750), Ricker
f0 = 1.25 Hz, dt = 0.035 s, 75 steps) is restated in tests/test_oracle3d.py
from the fixture's Mesh_Par_file, force.yaml and specfem_config.yaml.

Usage: python tests/golden/make_specfem3d_traces.py [/root/reference]
"""

import os
import sys

import numpy as np

BASE = "tests/unit-tests/displacement_tests/Newmark/serial/dim3"
CASE = "HomogeneousHalfspaceSmallNoABCForceSource"


def main(ref="/root/reference"):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    d = os.path.join(ref, BASE, CASE)
    names, nets, xs, ys, zs = [], [], [], [], []
    with open(os.path.join(d, "STATIONS")) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 6:
                names.append(p[0])
                nets.append(p[1])
                ys.append(float(p[2]))
                xs.append(float(p[3]))
                zs.append(float(p[5]))
    tdir = os.path.join(d, "traces")
    files = set(os.listdir(tdir))
    keep = [i for i, (n, net) in enumerate(zip(names, nets))
            if all(f"{net}.{n}.S3.BX{c}.sem{t}" in files for c in "XYZ" for t in "dva")]
    times = None
    traces = []
    for t in "dva":
        per_sta = []
        for i in keep:
            comps = []
            for c in "XYZ":
                arr = np.loadtxt(os.path.join(tdir, f"{nets[i]}.{names[i]}.S3.BX{c}.sem{t}"))
                times = arr[:, 0] if times is None else times
                comps.append(arr[:, 1])
            per_sta.append(comps)
        traces.append(per_sta)
    np.savez_compressed(os.path.join(out_dir, f"specfem3d_{CASE}.npz"),
                        names=np.array([names[i] for i in keep]),
                        x=np.array([xs[i] for i in keep]), y=np.array([ys[i] for i in keep]),
                        z=np.array([zs[i] for i in keep]), times=times,
                        types=np.array(list("dva")), traces=np.array(traces))
    print(f"{CASE}: {len(keep)} stations, {len(times)} samples")


if __name__ == "__main__":
    main(*sys.argv[1:])
