"""Sources, receivers and the seeded benchmark configurations.

Sources are vertical/oblique point forces with a Ricker or Dirac time
function, the reference's ``force`` source
(core/specfem/source/dim2/vector_source/force_source.cpp:22-73,
core/specfem/source_time_functions/ricker.cpp:9-66, dirac.cpp:7-52).  Receivers are station
coordinates (core/specfem/receivers/dim2/receiver.hpp).  The seeded
configurations follow the measured workloads: an 80x60 box of 50 m quads
(single homogeneous medium, free top, Stacey elsewhere) and an 800x600 box with
six horizontal layers.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh2D, box_mesh, elastic_moduli, gll_coordinates


STF_RICKER, STF_DIRAC, STF_HEAVISIDE = 0, 1, 2
SRC_FORCE, SRC_MOMENT = 0, 1


@dataclass
class Sources2D:
    x: np.ndarray
    z: np.ndarray
    angle: np.ndarray    # degrees, clockwise from +z (0 = vertical force)
    f0: np.ndarray
    tshift: np.ndarray
    factor: np.ndarray
    # time function per source: STF_RICKER, STF_DIRAC or STF_HEAVISIDE (None:
    # all Ricker); for a Heaviside source f0 holds the half duration hdur (s)
    stf: np.ndarray = None
    # source kind per source: SRC_FORCE or SRC_MOMENT (None: all forces), and
    # the moment tensors [n][3] = (Mxx, Mzz, Mxz) of the SRC_MOMENT sources
    kind: np.ndarray = None
    moment: np.ndarray = None

    def kinds(self):
        """Time-function kind per source."""
        n = len(self.x)
        return (np.zeros(n, np.int32) if self.stf is None
                else np.asarray(self.stf, dtype=np.int32).reshape(n))

    def source_kinds(self):
        """SRC_FORCE / SRC_MOMENT per source."""
        n = len(self.x)
        return (np.zeros(n, np.int32) if self.kind is None
                else np.asarray(self.kind, dtype=np.int32).reshape(n))

    def moments(self):
        n = len(self.x)
        return (np.zeros((n, 3), np.float32) if self.moment is None
                else np.asarray(self.moment, dtype=np.float32).reshape(n, 3))

    @staticmethod
    def moment_tensor(x, z, Mxx, Mzz, Mxz, f0=None, factor=1.0, tshift=0.0,
                      hdur=None):
        """Moment-tensor sources (core/specfem/source/dim2/tensor_source/
        moment_tensor_source.cpp) with a Ricker time function of dominant
        frequency f0, or a Heaviside (error-function step) of half duration
        hdur (core/specfem/source_time_functions/heaviside.cpp)."""
        if (f0 is None) == (hdur is None):
            raise ValueError("give exactly one of f0 (Ricker) and hdur (Heaviside)")
        s = Sources2D.ricker(x, z, f0=hdur if f0 is None else f0, factor=factor,
                             tshift=tshift)
        n = len(s.x)
        if hdur is not None:
            s.stf = np.full(n, STF_HEAVISIDE, np.int32)
        s.kind = np.full(n, SRC_MOMENT, np.int32)
        m = np.stack([np.broadcast_to(np.asarray(v, np.float32), (n,))
                      for v in (Mxx, Mzz, Mxz)], axis=1)
        s.moment = np.ascontiguousarray(m)
        return s

    @staticmethod
    def dirac(x, z, factor=1e10, angle=0.0, tshift=0.0):
        """Dirac sources: the reference approximates the impulse by a
        Gaussian of f0 = 1 / (10 dt) (dirac.cpp:23-35), so f0 is set from dt
        when the context is built (the f0 given here is ignored)."""
        s = Sources2D.ricker(x, z, f0=1.0, factor=factor, angle=angle, tshift=tshift)
        s.stf = np.full(len(s.x), STF_DIRAC, np.int32)
        return s

    @staticmethod
    def ricker(x, z, f0, factor=1e10, angle=0.0, tshift=0.0):
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        n = len(x)

        def full(v, dt):
            v = np.asarray(v, dtype=dt)
            return np.broadcast_to(v, (n,)).copy()

        return Sources2D(x=x, z=full(z, np.float64), angle=full(angle, np.float32),
                         f0=full(f0, np.float32), tshift=full(tshift, np.float32),
                         factor=full(factor, np.float32))

    def __len__(self):
        return len(self.x)


@dataclass
class Receivers2D:
    x: np.ndarray
    z: np.ndarray
    names: tuple = ()

    def __len__(self):
        return len(self.x)


@dataclass
class Config2D:
    name: str
    mesh: Mesh2D
    sources: Sources2D
    receivers: Receivers2D
    dt: float
    nstep: int
    nsbs: int = 1
    seismo_types: tuple = ("displacement",)
    meta: dict = None


def homogeneous_box(nx=80, nz=60, h=50.0, nstep=2000, dt=5e-4, nrec=50) -> Config2D:
    """4000x3000 m box (80x60 quads of 50 m), rho=2700, vp=3000, vs=1732, free
    top, Stacey on the other three sides; vertical Ricker force (f0=10 Hz,
    factor 1e10) at (xmax/2, 2 zmax/3) = (2000 m, 2000 m); receivers 100 m below the top,
    x = 40 + 80 i (chosen spacing)."""
    xmax, zmax = nx * h, nz * h
    mesh = box_mesh(0.0, xmax, 0.0, zmax, nx, nz, ngnod=4,
                    absorbing=("bottom", "right", "left"),
                    rho=2700.0, vp=3000.0, vs=1732.0)
    src = Sources2D.ricker([0.5 * xmax], [zmax * 2.0 / 3.0], f0=10.0, factor=1e10)
    rx = 40.0 + (xmax - 80.0) / max(nrec - 1, 1) * np.arange(nrec)
    rec = Receivers2D(x=rx, z=np.full(nrec, zmax - 100.0))
    return Config2D("homogeneous_80x60", mesh, src, rec, dt, nstep,
                    meta={"nx": nx, "nz": nz, "h": h})


def layered_box(nx=800, nz=600, h=50.0, nstep=5000, dt=4e-4, nsrc=8, nrec=500,
                nlayers=6, seed=1001, f0=5.0, acoustic=False) -> Config2D:
    """Layered box: ``nlayers`` equal horizontal layers with vp ~ U[2000, 6000]
    m/s, vs = vp/sqrt(3), rho = 310 vp^0.25 (Gardner), ``nsrc`` vertical Ricker
    sources (f0 = 5 Hz) at seeded interior positions and ``nrec`` surface
    receivers; free top, Stacey on the other three sides.  ``acoustic``: the
    same layers as fluids (vs = 0) with an acoustic free surface on top."""
    rng = np.random.default_rng(seed)
    vps = rng.uniform(2000.0, 6000.0, size=nlayers)
    xmax, zmax = nx * h, nz * h
    per = max(nz // nlayers, 1)

    def layer(jz):  # layer 0 at the top; jz may be an index array
        return np.minimum((nz - 1 - np.asarray(jz)) // per, nlayers - 1)

    mesh = box_mesh(0.0, xmax, 0.0, zmax, nx, nz, ngnod=4,
                    absorbing=("bottom", "right", "left"),
                    rho=lambda jx, jz: 310.0 * vps[layer(jz)] ** 0.25,
                    vp=lambda jx, jz: vps[layer(jz)],
                    vs=(0.0 if acoustic else lambda jx, jz: vps[layer(jz)] / np.sqrt(3.0)),
                    free=("top",) if acoustic else ())
    sx = rng.uniform(0.05 * xmax, 0.95 * xmax, size=nsrc)
    sz = rng.uniform(0.067 * zmax, 0.933 * zmax, size=nsrc)
    src = Sources2D.ricker(sx, sz, f0=f0, factor=1e10)
    rx = 40.0 + (xmax - 80.0) / max(nrec - 1, 1) * np.arange(nrec)
    rec = Receivers2D(x=rx, z=np.full(nrec, zmax))
    return Config2D(f"layered_{'acoustic_' if acoustic else ''}{nx}x{nz}", mesh, src, rec,
                    dt, nstep, meta={"nx": nx, "nz": nz, "h": h, "vp_layers": vps.tolist(),
                                     "seed": seed})


def gaussian_vs_anomaly(mesh: Mesh2D, x0=2000.0, z0=1500.0, sigma=300.0, amp=0.05,
                        rho=2700.0, vp=3000.0, vs=1732.0):
    """(kappa, mu, rho) at every GLL point of a homogeneous medium whose vs
    carries a Gaussian anomaly vs (1 + amp exp(-r^2 / (2 sigma^2)))."""
    xz = gll_coordinates(mesh)
    r2 = (xz[..., 0] - x0) ** 2 + (xz[..., 1] - z0) ** 2
    vs_pt = vs * (1.0 + amp * np.exp(-r2 / (2.0 * sigma * sigma)))
    shape = vs_pt.shape
    return elastic_moduli(np.full(shape, rho, np.float32), np.full(shape, vp, np.float32),
                          vs_pt.astype(np.float32))


def adjoint_case(nx=80, nz=60, nstep=2000, dt=5e-4, nrec=50):
    """Adjoint-gradient workload: the 80x60 homogeneous box, its source and
    50 receivers; the true model adds a 5% Gaussian vs anomaly (sigma 300 m)
    at (2000 m, 1500 m).  Returns (config with the homogeneous starting
    model, true (kappa, mu, rho))."""
    cfg = homogeneous_box(nx=nx, nz=nz, nstep=nstep, dt=dt, nrec=nrec)
    h = cfg.meta["h"]
    true = gaussian_vs_anomaly(cfg.mesh, x0=0.5 * nx * h, z0=0.5 * nz * h)
    cfg.name = f"adjoint_{nx}x{nz}"
    return cfg, true


def smooth_perturbation(mesh: Mesh2D, rng, nmodes=6, kmax=4):
    """A seeded smooth field at every GLL point (nspec, 5, 5) with unit RMS:
    a random combination of ``nmodes`` products sin(p pi x / Lx + a)
    cos(q pi z / Lz + b), 1 <= p, q <= kmax (finite-difference directions)."""
    xz = gll_coordinates(mesh)
    x0, x1 = float(xz[..., 0].min()), float(xz[..., 0].max())
    z0, z1 = float(xz[..., 1].min()), float(xz[..., 1].max())
    X = (xz[..., 0] - x0) / max(x1 - x0, 1e-30)
    Z = (xz[..., 1] - z0) / max(z1 - z0, 1e-30)
    f = np.zeros(X.shape)
    for _ in range(nmodes):
        p, q = rng.integers(1, kmax + 1, size=2)
        a, b = rng.uniform(0, 2 * np.pi, size=2)
        f += rng.standard_normal() * np.sin(p * np.pi * X + a) * np.cos(q * np.pi * Z + b)
    return f / np.sqrt(np.mean(f * f))


def batched_shots(nshots=64, seed=1003, nstep=3000, dt=4e-4, nrec=100, nmin=40, nmax=400,
                  h=50.0):
    """Independent shots on meshes of uneven size: shot k has nx ~ U[nmin, nmax]
    quads, nz = round(0.75 nx), six seeded layers (as layered_box), one Ricker
    source with f0 ~ U[5, 15] Hz at a seeded interior position and ``nrec``
    surface receivers; free top, Stacey elsewhere."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nshots):
        nx = int(rng.integers(nmin, nmax + 1))
        nz = int(round(0.75 * nx))
        f0 = float(rng.uniform(5.0, 15.0))
        cfg = layered_box(nx=nx, nz=nz, h=h, nstep=nstep, dt=dt, nsrc=1, nrec=nrec,
                          seed=int(rng.integers(1 << 30)), f0=f0)
        cfg.name = f"shot{k}_{nx}x{nz}"
        cfg.meta["f0"] = f0
        out.append(cfg)
    return out
