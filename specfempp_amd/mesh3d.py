"""Host model for the 3D elastic path: hex8 meshes, sources, receivers and the
seeded 3D workloads.

A mesh is what the reference reads from a 3D mesh database
(src/io/mesh/impl/fortran/dim3/mesh.cpp): control-node coordinates, per-element
8-node connectivity (node order of
core/specfem/shape_function/dim3/shape_functions3d.cpp: the z- face
counter-clockwise from (-,-,-), then the z+ face), material per GLL point and
the absorbing faces (include/enumerations/dim3/mesh_entities.hpp numbering:
bottom=1 (z-), right=2 (x+), top=3 (z+), left=4 (x-), front=5 (y-),
back=6 (y+)).  The structured box generator replaces the Fortran mesher for
the rectangular domains this path is measured on.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import elastic_moduli, gll_points

NGLL = 5
FACES = {"bottom": 1, "right": 2, "top": 3, "left": 4, "front": 5, "back": 6}


@dataclass
class Mesh3D:
    coord: np.ndarray      # (nnodes, 3) float64
    knods: np.ndarray      # (nspec, 8) int32
    kappa: np.ndarray      # (nspec, 5, 5, 5) float32, [e, iz, iy, ix]
    mu: np.ndarray
    rho: np.ndarray
    abs_ispec: np.ndarray  # (nabs,) int32
    abs_face: np.ndarray   # (nabs,) int32
    shape: tuple = (0, 0, 0)   # (nx, ny, nz) of a structured box
    meta: dict = field(default_factory=dict)

    @property
    def nspec(self) -> int:
        return int(self.knods.shape[0])

    @property
    def nnodes(self) -> int:
        return int(self.coord.shape[0])

    def validate(self):
        n = self.nspec
        if self.knods.shape != (n, 8):
            raise ValueError("knods must be (nspec, 8)")
        for name in ("kappa", "mu", "rho"):
            if getattr(self, name).shape != (n, NGLL, NGLL, NGLL):
                raise ValueError(f"{name} must be (nspec, 5, 5, 5)")
        return self


@dataclass
class Sources3D:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    kind: np.ndarray      # 0 force, 1 moment tensor
    comp: np.ndarray      # (n, 6): force (fx, fy, fz, 0, 0, 0) or (Mxx, Myy, Mzz, Mxy, Mxz, Myz)
    f0: np.ndarray
    tshift: np.ndarray
    factor: np.ndarray

    @staticmethod
    def make(x, y, z, *, force=None, moment=None, f0=1.0, tshift=0.0, factor=1.0):
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        n = len(x)
        if (force is None) == (moment is None):
            raise ValueError("give exactly one of force= or moment=")
        comp = np.zeros((n, 6), dtype=np.float32)
        if force is not None:
            comp[:, :3] = np.broadcast_to(np.asarray(force, dtype=np.float32), (n, 3))
            kind = np.zeros(n, np.int32)
        else:
            comp[:] = np.broadcast_to(np.asarray(moment, dtype=np.float32), (n, 6))
            kind = np.ones(n, np.int32)

        def full(v, dt):
            return np.broadcast_to(np.asarray(v, dtype=dt), (n,)).copy()

        return Sources3D(x=x, y=full(y, np.float64), z=full(z, np.float64), kind=kind,
                         comp=comp, f0=full(f0, np.float32), tshift=full(tshift, np.float32),
                         factor=full(factor, np.float32))

    def __len__(self):
        return len(self.x)


@dataclass
class Receivers3D:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    names: tuple = ()

    def __len__(self):
        return len(self.x)


@dataclass
class Config3D:
    name: str
    mesh: Mesh3D
    sources: Sources3D
    receivers: Receivers3D
    dt: float
    nstep: int
    nsbs: int = 1
    seismo_types: tuple = ("displacement",)
    meta: dict = None


def box_mesh3d(x0, x1, y0, y1, z0, z1, nx, ny, nz, *,
               absorbing=("bottom", "right", "left", "front", "back"),
               rho=2700.0, vp=3000.0, vs=1732.0) -> Mesh3D:
    """Structured nx x ny x nz hex8 box, elements ordered x fastest
    (ispec = (jz ny + jy) nx + jx).  rho/vp/vs are scalars, callables of the
    GLL point coordinates (x, y, z arrays of shape (nspec, 5, 5, 5)), or
    arrays of that shape."""
    nxn, nyn, nzn = nx + 1, ny + 1, nz + 1
    xs, ys, zs = (np.linspace(a, b, n) for a, b, n in ((x0, x1, nxn), (y0, y1, nyn),
                                                         (z0, z1, nzn)))
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coord = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    jz, jy, jx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = ((jz * nyn + jy) * nxn + jx).ravel()
    up = nyn * nxn
    knods = np.stack([b, b + 1, b + 1 + nxn, b + nxn,
                      b + up, b + up + 1, b + up + 1 + nxn, b + up + nxn], axis=1).astype(np.int32)
    mesh = Mesh3D(coord=coord, knods=knods, kappa=None, mu=None, rho=None,
                  abs_ispec=None, abs_face=None, shape=(nx, ny, nz))
    xyz = gll_coordinates3d(mesh)

    def field_of(val):
        if callable(val):
            return np.asarray(val(xyz[..., 0], xyz[..., 1], xyz[..., 2]), dtype=np.float64)
        return np.broadcast_to(np.asarray(val, dtype=np.float64), xyz.shape[:-1])

    mesh.kappa, mesh.mu, mesh.rho = elastic_moduli(field_of(rho).astype(np.float32),
                                                   field_of(vp).astype(np.float32),
                                                   field_of(vs).astype(np.float32))
    e = np.arange(nx * ny * nz)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    on = {1: ez == 0, 3: ez == nz - 1, 4: ex == 0, 2: ex == nx - 1, 5: ey == 0, 6: ey == ny - 1}
    parts = [(e[on[FACES[f]]], np.full(int(on[FACES[f]].sum()), FACES[f])) for f in absorbing]
    if parts:
        ai = np.concatenate([p[0] for p in parts])
        af = np.concatenate([p[1] for p in parts])
        key = np.lexsort((af, ai))
        ai, af = ai[key], af[key]
    else:
        ai = af = np.zeros(0, np.int64)
    mesh.abs_ispec, mesh.abs_face = ai.astype(np.int32), af.astype(np.int32)
    return mesh.validate()


def gll_coordinates3d(mesh: Mesh3D) -> np.ndarray:
    """(nspec, 5, 5, 5, 3) coordinates of every GLL point (float64)."""
    xi = gll_points()
    c = mesh.coord[mesh.knods]  # (nspec, 8, 3)
    G, E, X = np.meshgrid(xi, xi, xi, indexing="ij")  # [iz, iy, ix]
    N = np.stack([(1 - X) * (1 - E) * (1 - G), (1 + X) * (1 - E) * (1 - G),
                  (1 + X) * (1 + E) * (1 - G), (1 - X) * (1 + E) * (1 - G),
                  (1 - X) * (1 - E) * (1 + G), (1 + X) * (1 - E) * (1 + G),
                  (1 + X) * (1 + E) * (1 + G), (1 - X) * (1 + E) * (1 + G)], axis=-1) * 0.125
    return np.einsum("zyxn,end->ezyxd", N, c)


def halfspace_fixture():
    """The reference's 3D displacement-test set-up
    (HomogeneousHalfspaceSmallNoABCForceSource): 8000 x 7000 x 7500 m box of
    8 x 6 x 6 hex8 elements, rho 2700, vp 3000, vs 1732.051, free surfaces,
    vertical force 1e14 at (4000, 3500, -3750), Ricker f0 = 1.25 Hz, dt =
    0.035 s, 75 steps.  Receivers come from the fixture's station table."""
    mesh = box_mesh3d(0.0, 8000.0, 0.0, 7000.0, -7500.0, 0.0, 8, 6, 6, absorbing=(),
                      rho=2700.0, vp=3000.0, vs=1732.051)
    src = Sources3D.make([4000.0], [3500.0], [-3750.0], force=(0.0, 0.0, 1e14), f0=1.25)
    return mesh, src


def gaussian_random_field(shape_xyz, spacing, corr, rng):
    """Zero-mean, unit-std Gaussian random field on a regular grid, filtered
    in Fourier space with a Gaussian kernel of correlation length `corr`."""
    w = rng.standard_normal(shape_xyz)
    k = [np.fft.fftfreq(n, d=spacing) * 2 * np.pi for n in shape_xyz]
    KX, KY, KZ = np.meshgrid(*k, indexing="ij")
    filt = np.exp(-0.25 * corr * corr * (KX ** 2 + KY ** 2 + KZ ** 2))
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    return f / f.std()


def hex_box_case(n=64, h=100.0, nstep=4000, dt=1.25e-3, nrec_side=16, seed=1004,
                 moment=(1e15, 1e15, 1e15, 0.0, 0.0, 0.0), f0=2.0, perturb=0.02,
                 corr=500.0):
    """3D workload: n^3 hex8 elements of h metres (6.4 km cube at n = 64);
    vp linear from 3000 m/s at the top to 6000 m/s at the base, vs = vp/sqrt(3),
    rho = 310 vp^0.25 (Gardner), times (1 + perturb * g) with g a seeded
    Gaussian random field of correlation length `corr` (sampled at the GLL
    points from a grid of spacing h/4); free top, Stacey on the other five
    faces; moment-tensor source (explosion by default), Ricker f0 = 2 Hz, at
    3 km depth in the centre; nrec_side^2 receivers on the surface grid.

    The time step is 1.25e-3 s, not the nominal 2e-3 s: at h = 100 m with
    vp up to ~6.5 km/s, 2e-3 s violates the explicit Newmark limit (a uniform
    6.6 km/s box of 100 m hexes is stable at 1.5e-3 s and diverges at
    1.75e-3 s)."""
    L = n * h
    rng = np.random.default_rng(seed)
    m = max(4 * n + 1, 2)
    g = gaussian_random_field((m, m, m), L / (m - 1), corr, rng)

    def sample(x, y, z):
        ix = np.clip(np.rint(x / L * (m - 1)).astype(int), 0, m - 1)
        iy = np.clip(np.rint(y / L * (m - 1)).astype(int), 0, m - 1)
        iz = np.clip(np.rint((z + L) / L * (m - 1)).astype(int), 0, m - 1)
        return g[ix, iy, iz]

    def vp(x, y, z):
        depth = -z
        return (3000.0 + 3000.0 * depth / L) * (1.0 + perturb * sample(x, y, z))

    def vs(x, y, z):
        return vp(x, y, z) / np.sqrt(3.0)

    def rho(x, y, z):
        return 310.0 * vp(x, y, z) ** 0.25

    mesh = box_mesh3d(0.0, L, 0.0, L, -L, 0.0, n, n, n, rho=rho, vp=vp, vs=vs)
    src = Sources3D.make([L / 2], [L / 2], [-3000.0], moment=moment, f0=f0)
    r = (np.arange(nrec_side) + 0.5) * L / nrec_side
    RX, RY = np.meshgrid(r, r, indexing="ij")
    rec = Receivers3D(x=RX.ravel(), y=RY.ravel(), z=np.zeros(RX.size))
    return Config3D(f"hex{n}", mesh, src, rec, dt, nstep,
                    meta={"n": n, "h": h, "seed": seed, "perturb": perturb, "corr": corr})
