"""Reader of the 2D mesh database the reference's mesher writes
(xmeshfem2D's Fortran-unformatted ``database.bin``), producing a ``Mesh2D``.

The file is a sequence of Fortran sequential-unformatted records: a 4-byte
length, the payload, the same length again (include/io/fortranio/
fortran_io.tpp:43-65; int 4 B, float 4 B, double 8 B, logical 4 B, strings
512 B, setup/specfem_setup.hpp.in:7).  The records are interpreted in the
order the reference reads them (src/io/mesh/impl/fortran/dim2/mesh.cpp:26-215):

1. header: 64 records of run parameters, one of them optional
   (read_mesh_database.cpp:10-150); only nspec (record 3) and npgeo
   (record 4) are used;
2. control-node coordinates, npgeo records (ipoin, x, z)
   (read_mesh_database.cpp:152-178);
3. mesh parameters (numat, ngnod, nspec, pointsdisp, plot flag) and the
   boundary counts (read_parameters.cpp:10-60);
4. attenuation (n_sls, f0, flag) (read_mesh_database.cpp:180-204);
5. materials, numat records (n, indic, 13 doubles); isotropic materials
   are indic 1 (density, cp, cs): elastic with a non-zero shear speed,
   acoustic with cs = 0 (read_material_properties.cpp:160-250);
6. per element (n, material, ngnod control nodes, pml flag)
   (read_material_properties.cpp:360-410);
7. (ninterfaces, max_interface_size), the absorbing edges in database order
   (element, bottom/right/top/left flags, type, 8 ints)
   (read_boundaries.cpp:150-200), the acoustic free surface (element, 1 for a
   corner node or 2 for an edge, two control-node ids; the entity type comes
   from comparing those ids with the element's corner nodes,
   read_boundaries.cpp:236-300), the fluid-solid edges (acoustic element,
   elastic element; mesh.cpp:112-127, kept in ``info``: the coupling is
   rebuilt from the elements' shared edges), then tangential nodes and the
   adjacency graph, which are skipped.  Acoustic forcing, poroelastic
   interfaces and axial elements are rejected.

The reference adds corner entries to the absorbing list only when one
element carries both a horizontal and a vertical absorbing edge; its test
(read_boundaries.cpp:17-40) compares an edge's type with itself, so no corner
is ever added: the list is the database's edges, in order, as here.
"""

from __future__ import annotations

import gzip
import struct

import numpy as np

from .mesh import Mesh2D, acoustic_moduli, elastic_moduli

NGLL = 5
# header records after the optional one (AXISYM ... ADD_RANDOM_PERTURBATION)
_HEADER_HEAD = 42
_HEADER_TAIL = 22


def fortran_records(data: bytes):
    """Splits Fortran sequential-unformatted data into record payloads."""
    recs, pos, n = [], 0, len(data)
    while pos < n:
        if pos + 4 > n:
            raise ValueError("truncated record marker")
        (length,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + length
        if length < 0 or end + 4 > n:
            raise ValueError(f"bad record length {length} at byte {pos}")
        (length2,) = struct.unpack_from("<i", data, end)
        if length2 != length:
            raise ValueError(f"record markers disagree at byte {pos}")
        recs.append(data[pos + 4:end])
        pos = end + 4
    return recs


class _Cursor:
    def __init__(self, recs):
        self.recs, self.i = recs, 0

    def next(self, fmt=None):
        if self.i >= len(self.recs):
            raise ValueError("the database ends early")
        r = self.recs[self.i]
        self.i += 1
        if fmt is None:
            return r
        if struct.calcsize("<" + fmt) != len(r):
            raise ValueError(f"record {self.i - 1} has {len(r)} bytes, expected layout {fmt}")
        return struct.unpack("<" + fmt, r)


def read_mesh_database_2d(path_or_bytes):
    """Reads a 2D ``database.bin`` (optionally gzip-compressed) into a
    ``Mesh2D`` with per-GLL-point (kappa, mu, rho) from each element's
    material; returns (mesh, info) where info holds the database's sizes.
    Isotropic elastic and acoustic materials are accepted (mu = 0 marks an
    acoustic element)."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as fh:
            data = fh.read()
    if data[:2] == b"\x1f\x8b":
        data = gzip.decompress(data)
    cur = _Cursor(fortran_records(data))

    # 1. header
    hdr = [cur.next() for _ in range(_HEADER_HEAD)]
    (nspec,) = struct.unpack("<i", hdr[2])
    npgeo, nproc = struct.unpack("<ii", hdr[3])
    if len(cur.recs[cur.i]) == 16:  # ATTENUATION_PERMITTIVITY, ..., f0 (optional)
        cur.next()
    for _ in range(_HEADER_TAIL):
        cur.next()

    # 2. control nodes
    coord = np.zeros((npgeo, 2), dtype=np.float64)
    seen = np.zeros(npgeo, dtype=bool)
    for _ in range(npgeo):
        ip, x, z = cur.next("idd")
        if not 1 <= ip <= npgeo:
            raise ValueError("control-node index out of range")
        coord[ip - 1] = (x, z)
        seen[ip - 1] = True
    if not seen.all():
        raise ValueError("some control nodes have no coordinates")

    # 3. parameters
    numat, ngnod, nspec2, _pointsdisp, _plot = cur.next("iiiii")
    if nspec2 != nspec:
        raise ValueError("nspec differs between header and parameters")
    (nelemabs, nelem_acforcing, nelem_acoustic_surface, n_fs, n_fp, n_sp,
     n_tangential, n_axis) = cur.next("8i")

    # 4. attenuation
    cur.next("idi")

    # 5. materials
    mats = []
    for i in range(numat):
        vals = cur.next("ii13d")
        n, indic = vals[0], vals[1]
        if n != i + 1:
            raise ValueError("material sets out of order")
        v = vals[2:]
        if indic != 1:
            raise ValueError(f"material {n}: only isotropic materials are supported "
                             f"(indic {indic})")
        mats.append((v[0], v[1], v[2]))  # density, cp, cs (0: acoustic)

    # 6. elements
    knods = np.full((nspec, ngnod), -1, dtype=np.int32)
    kmat = np.zeros(nspec, dtype=np.int32)
    for _ in range(nspec):
        vals = cur.next(f"ii{ngnod}ii")
        n, km = vals[0], vals[1]
        if not 1 <= n <= nspec or not 1 <= km <= numat:
            raise ValueError("element or material index out of range")
        nodes = np.asarray(vals[2:2 + ngnod])
        if (nodes < 1).any() or (nodes > npgeo).any():
            raise ValueError("control-node id out of range")
        knods[n - 1] = nodes - 1
        kmat[n - 1] = km - 1
    if (knods < 0).any():
        raise ValueError("some elements were not listed")

    # 7. boundaries
    cur.next("ii")  # ninterfaces, max_interface_size
    abs_ispec = np.zeros(nelemabs, dtype=np.int32)
    abs_type = np.zeros(nelemabs, dtype=np.int32)
    for k in range(max(nelemabs, 0)):
        vals = cur.next("i4ii8i")
        num, codes = vals[0], vals[1:5]
        if not 1 <= num <= nspec:
            raise ValueError("absorbing element out of range")
        if sum(1 for c in codes if c) != 1:
            raise ValueError("an absorbing line must cite exactly one edge")
        abs_ispec[k] = num - 1
        abs_type[k] = 1 + [bool(c) for c in codes].index(True)  # bottom, right, top, left
    free_ispec = np.zeros(max(nelem_acoustic_surface, 0), dtype=np.int32)
    free_type = np.zeros(max(nelem_acoustic_surface, 0), dtype=np.int32)
    for k in range(max(nelem_acoustic_surface, 0)):
        num, kind, e1, e2 = cur.next("4i")
        if not 1 <= num <= nspec:
            raise ValueError("acoustic free-surface element out of range")
        c = knods[num - 1, :4]
        e1, e2 = e1 - 1, e2 - 1
        if kind == 1:  # a corner node: bottom-left, bottom-right, top-right, top-left
            hit = [t for t, node in zip((5, 6, 7, 8), c) if node == e1]
        else:          # an edge: bottom, left, right, top by its two corners
            pairs = ((1, c[0], c[1]), (4, c[0], c[3]), (2, c[1], c[2]), (3, c[2], c[3]))
            hit = [t for t, a, b in pairs if {e1, e2} == {a, b}]
        if not hit:
            raise ValueError("could not type an acoustic free-surface entity")
        free_ispec[k] = num - 1
        free_type[k] = hit[0]
    if nelem_acforcing or n_fp or n_sp or n_axis:
        raise ValueError("acoustic forcing, poroelastic or axisymmetric meshes are not "
                         "supported")
    # fluid-solid edges (mesh.cpp:112-127): (acoustic element, elastic element)
    # pairs; the coupling itself is rebuilt from the shared edges
    fluid_solid = []
    for _ in range(max(n_fs, 0)):
        a, b = cur.next("ii")
        if not (1 <= a <= nspec and 1 <= b <= nspec):
            raise ValueError("fluid-solid element out of range")
        fluid_solid.append((a - 1, b - 1))
    cur.next("ii")  # tangential-curve flags
    for _ in range(n_tangential):
        cur.next("dd")
    (nadj,) = cur.next("i")
    for _ in range(nadj):
        cur.next("4i")
    if cur.i != len(cur.recs):
        raise ValueError("records left after the adjacency graph")

    # material at every GLL point (element-constant), FP32 as the reference
    rho = np.array([m[0] for m in mats], dtype=np.float64)[kmat].astype(np.float32)
    vp = np.array([m[1] for m in mats], dtype=np.float64)[kmat].astype(np.float32)
    vs = np.array([m[2] for m in mats], dtype=np.float64)[kmat].astype(np.float32)
    shape = (nspec, NGLL, NGLL)
    rb, pb, sb = (np.broadcast_to(x[:, None, None], shape) for x in (rho, vp, vs))
    kappa, mu, rho_p = elastic_moduli(rb, pb, sb)
    fluid = sb == 0
    if fluid.any():
        ka, _, _ = acoustic_moduli(rb, pb)
        kappa = np.where(fluid, ka, kappa).astype(np.float32)
        mu = np.where(fluid, np.float32(0), mu).astype(np.float32)
    mesh = Mesh2D(coord=coord, knods=knods, kappa=kappa, mu=mu, rho=np.ascontiguousarray(rho_p),
                  abs_ispec=abs_ispec, abs_type=abs_type,
                  meta={"source": "database", "numat": numat, "nproc": nproc},
                  free_ispec=free_ispec, free_type=free_type)
    info = {"nspec": nspec, "npgeo": npgeo, "ngnod": ngnod, "numat": numat,
            "nelemabs": nelemabs, "nelem_acoustic_surface": nelem_acoustic_surface,
            "fluid_solid": fluid_solid,
            "materials": mats, "nadjacency": nadj}
    return mesh, info
