"""Host mesh model for the 2D elastic P-SV path.

A mesh is what the reference reads from an xmeshfem2D database
(src/io/mesh/impl/fortran/dim2/mesh.cpp:26-215): control-node coordinates,
per-element control-node connectivity (4- or 9-node quads, node order of
core/specfem/shape_function/dim2/shape_functions2d.cpp), material properties
and the list of absorbing edges (include/enumerations/dim2/mesh_entities.hpp:29-38
numbering: bottom=1, right=2, top=3, left=4).  The structured box generator
below replaces the Fortran mesher for the rectangular domains this path is
measured on.

A GLL point with zero shear modulus belongs to an acoustic (fluid) element:
the reference's material reader makes an isotropic material with cs = 0
acoustic (src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:178-204).
Acoustic free-surface (Dirichlet) entities are edges 1..4 or corners 5..8
(bottom-left, bottom-right, top-right, top-left; the mesh-entity types of
src/io/mesh/impl/fortran/dim2/read_boundaries.cpp:236-275).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NGLL = 5

BOTTOM, RIGHT, TOP, LEFT = 1, 2, 3, 4
BOTTOM_LEFT, BOTTOM_RIGHT, TOP_RIGHT, TOP_LEFT = 5, 6, 7, 8
EDGE_TYPES = {"bottom": BOTTOM, "right": RIGHT, "top": TOP, "left": LEFT}


def acoustic_moduli(rho, vp):
    """(kappa, mu = 0, rho) of a fluid in FP32: ``kappa = density*cp*cp`` in
    float (include/medium/dim2/acoustic/isotropic/material.hpp:49-53)."""
    rho = np.asarray(rho, dtype=np.float32)
    vp = np.asarray(vp, dtype=np.float32)
    kappa = (rho * vp) * vp
    return kappa.astype(np.float32), np.zeros_like(kappa, dtype=np.float32), rho.copy()


def elastic_moduli(rho, vp, vs):
    """(kappa, mu, rho) from (rho, vp, vs) in FP32.

    Follows the promotions of the reference material constructor
    (include/medium/dim2/elastic/isotropic/material.hpp:62-69):
    ``mu = rho*cs*cs`` in float, ``kappa = rho*(cp*cp - (4.0/3.0)*cs*cs)``
    with ``cp*cp`` in float and the rest in double, rounded to float.
    """
    rho = np.asarray(rho, dtype=np.float32)
    vp = np.asarray(vp, dtype=np.float32)
    vs = np.asarray(vs, dtype=np.float32)
    mu = (rho * vs) * vs  # float32 arithmetic
    cp2 = (vp * vp).astype(np.float64)
    kappa = (rho.astype(np.float64) *
             (cp2 - ((4.0 / 3.0) * vs.astype(np.float64)) * vs.astype(np.float64)))
    return kappa.astype(np.float32), mu.astype(np.float32), rho.copy()


@dataclass
class Mesh2D:
    coord: np.ndarray          # (nnodes, 2) float64, control-node (x, z)
    knods: np.ndarray          # (nspec, ngnod) int32
    kappa: np.ndarray          # (nspec, NGLL, NGLL) float32, [e, iz, ix]
    mu: np.ndarray
    rho: np.ndarray
    abs_ispec: np.ndarray      # (nabs,) int32
    abs_type: np.ndarray       # (nabs,) int32
    nx: int = 0                # structured-box metadata (0 if unstructured)
    nz: int = 0
    extent: tuple = (0.0, 0.0, 0.0, 0.0)   # xmin, xmax, zmin, zmax
    meta: dict = field(default_factory=dict)
    # acoustic free-surface entities (element, type 1..8), database order
    free_ispec: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    free_type: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    @property
    def acoustic(self) -> np.ndarray:
        """(nspec,) bool: elements whose every GLL point has mu == 0."""
        return (self.mu.reshape(self.nspec, -1) == 0).all(axis=1)

    @property
    def nspec(self) -> int:
        return int(self.knods.shape[0])

    @property
    def ngnod(self) -> int:
        return int(self.knods.shape[1])

    @property
    def nnodes(self) -> int:
        return int(self.coord.shape[0])

    def expected_nglob(self) -> int:
        if self.nx and self.nz:
            return (4 * self.nx + 1) * (4 * self.nz + 1)
        raise ValueError("nglob is only known in closed form for a box")

    def validate(self):
        ns = self.nspec
        if self.ngnod not in (4, 9):
            raise ValueError("ngnod must be 4 or 9")
        if self.knods.min() < 0 or self.knods.max() >= self.nnodes:
            raise ValueError("knods out of range")
        for name in ("kappa", "mu", "rho"):
            arr = getattr(self, name)
            if arr.shape != (ns, NGLL, NGLL) or arr.dtype != np.float32:
                raise ValueError(f"{name} must be float32 of shape {(ns, NGLL, NGLL)}")
        if self.abs_ispec.shape != self.abs_type.shape:
            raise ValueError("abs_ispec/abs_type size mismatch")
        if len(self.abs_ispec) and (self.abs_ispec.min() < 0 or self.abs_ispec.max() >= ns):
            raise ValueError("absorbing edge element out of range")
        if len(self.abs_type) and not np.isin(self.abs_type, [1, 2, 3, 4]).all():
            raise ValueError("absorbing edge type must be 1..4")
        self.free_ispec = np.asarray(self.free_ispec, dtype=np.int32)
        self.free_type = np.asarray(self.free_type, dtype=np.int32)
        if self.free_ispec.shape != self.free_type.shape:
            raise ValueError("free_ispec/free_type size mismatch")
        if len(self.free_ispec) and (self.free_ispec.min() < 0 or self.free_ispec.max() >= ns):
            raise ValueError("free-surface element out of range")
        if len(self.free_type) and not np.isin(self.free_type, range(1, 9)).all():
            raise ValueError("free-surface type must be 1..8")
        return self


def gll_points():
    s = np.sqrt(3.0 / 7.0)
    return np.array([-1.0, -s, 0.0, s, 1.0])


def box_mesh(xmin, xmax, zmin, zmax, nx, nz, *, ngnod=4, absorbing=("bottom", "right", "left"),
             rho=2700.0, vp=3000.0, vs=1732.0, free=()) -> Mesh2D:
    """Structured nx-by-nz box; elements row-major from the bottom-left
    (ispec = iz*nx + ix).  ``rho``, ``vp``, ``vs`` are scalars or callables
    f(ix_elem, iz_elem) -> value (per element); elements with vs == 0 are
    acoustic.  ``free`` names the sides that are acoustic free surfaces
    (listed per side, bottom/right/top/left, elements in order)."""
    if ngnod == 4:
        nxn, nzn = nx + 1, nz + 1
        xs = np.linspace(xmin, xmax, nxn)
        zs = np.linspace(zmin, zmax, nzn)
        X, Z = np.meshgrid(xs, zs)            # (nzn, nxn)
        coord = np.stack([X.ravel(), Z.ravel()], axis=1)
        iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
        n0 = (iz * nxn + ix).ravel()
        knods = np.stack([n0, n0 + 1, n0 + 1 + nxn, n0 + nxn], axis=1)
    elif ngnod == 9:
        nxn, nzn = 2 * nx + 1, 2 * nz + 1
        xs = np.linspace(xmin, xmax, nxn)
        zs = np.linspace(zmin, zmax, nzn)
        X, Z = np.meshgrid(xs, zs)
        coord = np.stack([X.ravel(), Z.ravel()], axis=1)
        iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
        b = ((2 * iz) * nxn + 2 * ix).ravel()
        knods = np.stack([b, b + 2, b + 2 + 2 * nxn, b + 2 * nxn,      # corners
                          b + 1, b + 2 + nxn, b + 1 + 2 * nxn, b + nxn,  # mid-sides
                          b + 1 + nxn], axis=1)                         # centre
    else:
        raise ValueError("ngnod must be 4 or 9")
    knods = knods.astype(np.int32)

    def per_elem(val):
        if callable(val):
            JZ, JX = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
            try:  # vectorised callables take index arrays
                out = np.broadcast_to(np.asarray(val(JX, JZ), dtype=np.float64), (nz, nx))
            except (TypeError, ValueError):
                out = np.empty((nz, nx), dtype=np.float64)
                for jz in range(nz):
                    for jx in range(nx):
                        out[jz, jx] = val(jx, jz)
            return np.ascontiguousarray(out).ravel()
        return np.full(nx * nz, float(val))

    r, p, s = per_elem(rho), per_elem(vp), per_elem(vs)
    rr = np.repeat(r.astype(np.float32), NGLL * NGLL).reshape(-1, NGLL, NGLL)
    pp = np.repeat(p.astype(np.float32), NGLL * NGLL).reshape(-1, NGLL, NGLL)
    ss = np.repeat(s.astype(np.float32), NGLL * NGLL).reshape(-1, NGLL, NGLL)
    kappa, mu, rho_pt = elastic_moduli(rr, pp, ss)
    fluid = ss == 0
    if fluid.any():
        ka, _, _ = acoustic_moduli(rr, pp)
        kappa = np.where(fluid, ka, kappa).astype(np.float32)
        mu = np.where(fluid, np.float32(0), mu).astype(np.float32)

    # Absorbing edges in element order; within an element bottom, right, top,
    # left (so at a corner GLL point the later edge's normal is kept, as in the
    # reference's last-edge-wins assembly).
    want = {EDGE_TYPES[a] for a in absorbing}
    ids = np.arange(nx * nz)
    jz_e, jx_e = ids // nx, ids % nx
    on = {BOTTOM: jz_e == 0, RIGHT: jx_e == nx - 1, TOP: jz_e == nz - 1, LEFT: jx_e == 0}
    order = {BOTTOM: 0, RIGHT: 1, TOP: 2, LEFT: 3}
    parts = [(ids[on[t]], np.full(int(on[t].sum()), t)) for t in (BOTTOM, RIGHT, TOP, LEFT)
             if t in want]
    if parts:
        ai = np.concatenate([p[0] for p in parts])
        at = np.concatenate([p[1] for p in parts])
        rank = np.array([order[t] for t in at.tolist()]) if len(at) else np.zeros(0, int)
        key = np.lexsort((rank, ai))  # element order, then bottom/right/top/left
        ai, at = ai[key], at[key]
    else:
        ai, at = np.zeros(0, np.int64), np.zeros(0, np.int64)
    fparts = [(ids[on[EDGE_TYPES[a]]], np.full(int(on[EDGE_TYPES[a]].sum()), EDGE_TYPES[a]))
              for a in ("bottom", "right", "top", "left") if a in free]
    fi = np.concatenate([p[0] for p in fparts]) if fparts else np.zeros(0, np.int64)
    ft = np.concatenate([p[1] for p in fparts]) if fparts else np.zeros(0, np.int64)
    return Mesh2D(coord=coord.astype(np.float64), knods=knods, kappa=kappa, mu=mu,
                  rho=rho_pt, abs_ispec=ai.astype(np.int32),
                  abs_type=at.astype(np.int32), nx=nx, nz=nz,
                  extent=(xmin, xmax, zmin, zmax), free_ispec=fi.astype(np.int32),
                  free_type=ft.astype(np.int32)).validate()


def gll_coordinates(mesh: Mesh2D) -> np.ndarray:
    """(nspec, NGLL, NGLL, 2) physical coordinates of every GLL point (float64,
    only for set-up and plotting; the solver computes its own in FP32)."""
    xi = gll_points()
    c = mesh.coord[mesh.knods]  # (nspec, ngnod, 2)
    XI, GA = np.meshgrid(xi, xi)  # [iz, ix]
    if mesh.ngnod == 4:
        N = np.stack([0.25 * (XI - 1) * (GA - 1), -0.25 * (XI + 1) * (GA - 1),
                      0.25 * (XI + 1) * (GA + 1), -0.25 * (XI - 1) * (GA + 1)], axis=-1)
    else:
        N = np.stack([0.25 * (XI - 1) * XI * (GA - 1) * GA, 0.25 * (XI + 1) * XI * (GA - 1) * GA,
                      0.25 * (XI + 1) * XI * (GA + 1) * GA, 0.25 * (XI - 1) * XI * (GA + 1) * GA,
                      0.5 * (1 - XI * XI) * (GA - 1) * GA, 0.5 * (XI + 1) * XI * (1 - GA * GA),
                      0.5 * (1 - XI * XI) * (GA + 1) * GA, 0.5 * (XI - 1) * XI * (1 - GA * GA),
                      (1 - XI * XI) * (1 - GA * GA)], axis=-1)
    return np.einsum("zxn,end->ezxd", N, c)
