"""The 2D mesh-database reader on the reference's own database.bin fixtures
(copied, compressed, by tests/golden/make_specfem2d_traces.py): the mesh it
returns is the 80x60 Q9 box the displacement tests describe, its absorbing
edges are the database's in order, and the oracle run on it reproduces the
SPECFEM2D traces with the reference's criterion."""

import gzip
import os

import numpy as np
import pytest

from specfempp_amd.meshdb import fortran_records, read_mesh_database_2d
from tests._cases import FIXTURES, GOLDEN, fixture_case, reference_metric, rel_l2


def _db(name):
    return os.path.join(GOLDEN, f"specfem2d_{name}.database.bin.gz")


def _element_nodes(mesh):
    """Per element: sorted tuple of its control-node coordinates."""
    xz = mesh.coord[mesh.knods]  # (nspec, ngnod, 2)
    return {tuple(sorted(map(tuple, e))) for e in np.round(xz, 6)}


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_database_is_the_described_box(name):
    mesh, info = read_mesh_database_2d(_db(name))
    mesh.validate()
    assert (info["nspec"], info["npgeo"], info["ngnod"], info["numat"]) == (4800, 19481, 9, 1)
    assert info["materials"] == [(2700.0, 3000.0, 1732.051)]
    ref, _ = fixture_case(name)
    # same elements (as sets of control-node coordinates), same material
    assert _element_nodes(mesh) == _element_nodes(ref.mesh)
    for a, b in ((mesh.kappa, ref.mesh.kappa), (mesh.mu, ref.mesh.mu), (mesh.rho, ref.mesh.rho)):
        assert np.array_equal(np.unique(a), np.unique(b))
    # Q9 node order: corners, then edge midpoints, then the centre
    xz = mesh.coord[mesh.knods]
    np.testing.assert_allclose(xz[:, 8], xz[:, :4].mean(axis=1), atol=1e-9)
    np.testing.assert_allclose(xz[:, 4], 0.5 * (xz[:, 0] + xz[:, 1]), atol=1e-9)


def test_database_absorbing_edges_in_database_order():
    mesh, info = read_mesh_database_2d(_db("HomogeneousElasticDomainStaceyDirichletBC"))
    assert info["nelemabs"] == 280
    assert np.bincount(mesh.abs_type, minlength=5).tolist() == [0, 80, 60, 80, 60]
    # every cited edge lies on its side of the box (1 bottom, 2 right, 3 top, 4 left)
    corners = mesh.coord[mesh.knods[mesh.abs_ispec, :4]]  # (nabs, 4, 2)
    side = {1: (0, 1, 1, 0.0), 2: (1, 2, 0, 4000.0), 3: (2, 3, 1, 3000.0), 4: (3, 0, 0, 0.0)}
    for k, t in enumerate(mesh.abs_type):
        a, b, axis, val = side[int(t)]
        assert abs(corners[k, a, axis] - val) < 1e-6 and abs(corners[k, b, axis] - val) < 1e-6
    # the corner element cites two edges, in database order
    assert mesh.abs_ispec[0] == mesh.abs_ispec[1] == 0
    assert mesh.abs_type[:2].tolist() == [1, 4]


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_on_database_matches_specfem2d(oracle_lib, name):
    cfg, g = fixture_case(name)
    mesh, _ = read_mesh_database_2d(_db(name))
    o = oracle_lib.Oracle2D(mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    s = o.run(types=cfg.seismo_types)
    err = reference_metric(s, g["traces"])
    restated = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt,
                                   cfg.nstep).run(types=cfg.seismo_types)
    print(f"{name}: database mesh vs SPECFEM2D {err:.2e}; vs restated box "
          f"{rel_l2(s, restated):.2e}")
    assert err < 1e-3
    assert rel_l2(s, restated) < 1e-5


def test_reader_rejects_damaged_files():
    data = gzip.decompress(open(_db("HomogeneousElasticDomain"), "rb").read())
    assert len(fortran_records(data)) > 20000
    with pytest.raises(ValueError):
        read_mesh_database_2d(data[:-100])          # truncated
    bad = bytearray(data)
    bad[4 + 512] ^= 0xFF                             # first record's closing marker
    with pytest.raises(ValueError):
        read_mesh_database_2d(bytes(bad))


ACOUSTIC = ("HomogeneousAcousticDomain", "HomogeneousAcousticDomainDirichletBC",
            "HomogeneousAcousticDomainCompositeStaceyDirichletBC")


@pytest.mark.parametrize("name", ACOUSTIC)
def test_acoustic_databases(name):
    """The acoustic fixtures: one acoustic material (cs = 0; kappa =
    density cp^2 in float), free-surface edges typed from their control
    nodes, each lying on its side of the box."""
    mesh, info = read_mesh_database_2d(_db(name))
    mesh.validate()
    assert info["materials"] == [(2700.0, 3000.0, 0.0)]
    assert mesh.acoustic.all()
    assert np.unique(mesh.kappa).tolist() == [np.float32(np.float32(2700.0 * 3000.0) * 3000.0)]
    composite = name.endswith("CompositeStaceyDirichletBC")
    want = [0, 0, 80, 0] if composite else [80, 60, 80, 60]
    assert np.bincount(mesh.free_type, minlength=9)[1:5].tolist() == want
    assert len(mesh.abs_ispec) == (200 if composite else 0)
    corners = mesh.coord[mesh.knods[mesh.free_ispec, :4]]
    side = {1: (0, 1, 1, 0.0), 2: (1, 2, 0, 4000.0), 3: (2, 3, 1, 3000.0), 4: (3, 0, 0, 0.0)}
    for k, t in enumerate(mesh.free_type):
        a, b, axis, val = side[int(t)]
        assert abs(corners[k, a, axis] - val) < 1e-6 and abs(corners[k, b, axis] - val) < 1e-6
