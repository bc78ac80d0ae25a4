"""Problem set-ups shared by the oracle and GPU parity tests."""

import os

import numpy as np

from specfempp_amd.config import Config2D, Receivers2D, Sources2D, homogeneous_box, layered_box
from specfempp_amd.mesh import box_mesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# The two elastic SPECFEM2D fixtures of the reference's displacement tests
# (tests/unit-tests/displacement_tests/Newmark/serial/dim2/*/{Par_file,
# sources.yaml, specfem_config.yaml}): 80x60 nine-node quads on a 4000x3000 m
# box, rho=2700, vp=3000, vs=1732.051, dt=1.1e-3, 300 steps, Ricker f0=10 Hz,
# factor 1e10, vertical forces.
FIXTURES = {
    "HomogeneousElasticDomain": dict(
        absorbing=(), sources=[(2500.0, 2500.0)]),
    "HomogeneousElasticDomainStaceyDirichletBC": dict(
        absorbing=("bottom", "right", "top", "left"),
        sources=[(500.0, 500.0), (3500.0, 500.0), (500.0, 2500.0), (3500.0, 2500.0)]),
}
TYPE_NAMES = {"d": "displacement", "v": "velocity", "a": "acceleration"}


def fixture_case(name):
    g = np.load(os.path.join(GOLDEN, f"specfem2d_{name}.npz"))
    spec = FIXTURES[name]
    mesh = box_mesh(0.0, 4000.0, 0.0, 3000.0, 80, 60, ngnod=9, absorbing=spec["absorbing"],
                    rho=2700.0, vp=3000.0, vs=1732.051)
    src = Sources2D.ricker([s[0] for s in spec["sources"]], [s[1] for s in spec["sources"]],
                           f0=10.0, factor=1e10)
    rec = Receivers2D(x=g["x"], z=g["z"], names=tuple(g["names"]))
    types = tuple(TYPE_NAMES[t] for t in g["types"])
    cfg = Config2D(name, mesh, src, rec, 1.1e-3, 300, seismo_types=types)
    return cfg, g


def reference_metric(ours, traces):
    """Worst per-station sum|s - ref| / sum|ref| over components and types
    (tests/unit-tests/displacement_tests/Newmark/dim2/newmark_tests.cpp:337-362).
    ours: (nt, ntypes, nsta, 2); traces: (ntypes, nsta, 2, nt)."""
    worst = 0.0
    for it in range(traces.shape[0]):
        for ista in range(traces.shape[1]):
            a = np.asarray(ours[:, it, ista, :], dtype=np.float64).T
            b = traces[it, ista]
            worst = max(worst, float(np.abs(a - b).sum() / np.abs(b).sum()))
    return worst


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small_homogeneous(nstep=400):
    """The 80x60 homogeneous box (free top, Stacey elsewhere), 50 receivers."""
    return homogeneous_box(nstep=nstep)


def small_layered(nstep=300, nx=120, nz=90):
    """A reduced layered box (same generator as the 800x600 workload)."""
    return layered_box(nx=nx, nz=nz, nstep=nstep, nsrc=3, nrec=40, seed=1001)
