"""The host set-up (include/mxmesh.h): mesh contract, FP64 geometry, renumbering, the face2
stencil and its least-squares weights, checked against what the mathematics fixes."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes

FIXTURES = {"hex": 27, "tet-gmsh": 1101, "poly": 336, "hybrid": 216}


def fixture(name, order=None):
    return meshes.read_polymesh(cases.GOLDEN / "meshes" / name, order=order)


@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixture_sizes_and_unit_cube_volume(name):
    """tests/data/mesh/README.md: every fixture fills the unit cube."""
    m = fixture(name)
    assert m.n_cells == FIXTURES[name]
    assert abs(m.cell_volume.sum() - 1.0) <= 1e-13
    assert np.all(m.cell_volume > 0)


def all_meshes():
    yield "hex", fixture("hex")
    yield "tet-gmsh", fixture("tet-gmsh")
    yield "poly", fixture("poly")
    yield "hybrid", fixture("hybrid")
    yield "kuhn", meshes.kuhn_tets(4)
    yield "kuhn-morton", meshes.kuhn_tets(4, order="morton")
    yield "bump", meshes.bump_hex(12, 4, 3)
    yield "bump-hybrid", meshes.bump_hybrid(12, 10, 4)


@pytest.mark.parametrize("name,m", list(all_meshes()))
def test_mesh_contract(name, m):
    """core/include/autan/core/mesh.hpp:277-282: owner < neighbour, interior faces by (owner,
    neighbour), boundary faces by (zone, owner), each cell's faces ascending."""
    ni = m.n_interior
    own, nei = m.owner, m.neighbour
    assert np.all(own[:ni] < nei)
    key = own[:ni].astype(np.int64) * m.n_cells + nei
    assert np.all(np.diff(key) >= 0)
    bkey = m.face_zone.astype(np.int64) * m.n_cells + own[ni:]
    assert np.all(np.diff(bkey) >= 0)
    off, cf = m.cell_face_offsets, m.cell_faces
    for c in range(m.n_cells):
        faces = cf[off[c] : off[c + 1]]
        assert np.all(np.diff(faces) > 0)
        for f in faces:
            assert own[f] == c or (f < ni and nei[f] == c)


@pytest.mark.parametrize("name,m", list(all_meshes()))
def test_closed_cells_and_oriented_faces(name, m):
    """Each cell's outward area vectors sum to zero (a closed surface), normals are unit and
    point from the owner to the neighbour; r_left / r_right join the centroids through the face."""
    nf, ni, nc = m.n_faces, m.n_interior, m.n_cells
    n = m.face_normal.reshape(3, nf).T
    a = m.face_area
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, rtol=1e-15)
    s = np.zeros((nc, 3))
    np.add.at(s, m.owner, n * a[:, None])
    np.subtract.at(s, m.neighbour, n[:ni] * a[:ni, None])
    scale = np.zeros(nc)
    np.add.at(scale, m.owner, a)
    np.add.at(scale, m.neighbour, a[:ni])
    assert np.all(np.linalg.norm(s, axis=1) <= 1e-13 * scale)
    rl = m.r_left.reshape(3, nf).T
    rr = m.r_right.reshape(3, nf).T
    # owner to neighbour along the normal: x_nb - x_own = r_left - r_right
    assert np.all(np.einsum("ij,ij->i", rl[:ni] - rr[:ni], n[:ni]) > 0)
    assert np.all(np.einsum("ij,ij->i", rl[ni:], n[ni:]) > 0)
    if not m.all_planar:
        assert np.any(m.face_moment != 0)
    else:
        assert np.all(m.face_moment == 0)


def test_planarity_flags():
    assert fixture("hex").all_planar and fixture("tet-gmsh").all_planar and fixture("hybrid").all_planar
    assert not fixture("poly").all_planar
    assert meshes.kuhn_tets(3).all_planar and meshes.periodic_hex(5).all_planar


def test_bump_hybrid_cells():
    """Prisms on the wall, merged 2 x 2 x 2 polyhedra in the band, hexahedra elsewhere; the cells
    fill the channel less the bump, and the band's faces between merged cells are 8-node rings,
    non-planar over the bump."""
    nx, ny, nz = 12, 10, 4
    m = meshes.bump_hybrid(nx, ny, nz, prism_layers=2, band=(4, 8), order=None)
    prisms = 2 * nx * 2 * nz
    merged = (nx // 2) * 2 * (nz // 2)
    hexes = nx * (ny - 2 - 4) * nz
    assert m.n_cells == prisms + merged + hexes
    nfc = np.diff(m.cell_face_offsets)
    assert np.sum(nfc == 5) == prisms and np.sum(nfc == 6) == hexes and np.sum(nfc >= 9) == merged
    nodes = np.diff(m.face_node_offsets)
    assert set(np.unique(nodes)) == {3, 4, 8}
    # channel 3 x 1 x 1 less the bump's area, h w sqrt(pi) with h = 0.1, w = 0.2 (sampled)
    x = np.linspace(0.0, 3.0, nx + 1)
    wall = 0.1 * np.exp(-(((x - 1.5) / 0.2) ** 2))
    area = np.sum(0.5 * (wall[1:] + wall[:-1]) * np.diff(x))
    assert abs(m.cell_volume.sum() - (3.0 - area)) <= 1e-12
    assert not m.all_planar
    bad = [(5, 10, 4, 2, (4, 8)), (12, 10, 4, 2, (4, 7)), (12, 10, 4, 5, (4, 8))]
    for nx_, ny_, nz_, pl, band in bad:
        with pytest.raises(RuntimeError):
            meshes.bump_hybrid(nx_, ny_, nz_, prism_layers=pl, band=band)


def test_periodic_hex():
    n = 6
    m = meshes.periodic_hex(n, -10.0, 10.0)
    assert m.n_cells == n**3 and m.n_faces == 3 * n**3 and m.n_interior == m.n_faces
    assert np.all(np.diff(m.cell_face_offsets) == 6)
    np.testing.assert_allclose(m.cell_volume, (20.0 / n) ** 3, rtol=1e-14)
    d = (m.r_left - m.r_right).reshape(3, -1)
    # every face joins two images one spacing apart along its normal
    np.testing.assert_allclose(np.linalg.norm(d, axis=0), 20.0 / n, rtol=1e-13)
    assert np.all(np.diff(m.stencil_offsets) == 24)


def test_kuhn_tets_are_reproducible_and_keep_the_cube():
    a = meshes.kuhn_tets(5, 0.1, 0, order=None)
    b = meshes.kuhn_tets(5, 0.1, 0, order=None)
    c = meshes.kuhn_tets(5, 0.1, 1, order=None)
    assert np.array_equal(a.x, b.x) and not np.array_equal(a.x, c.x)
    assert a.n_cells == 6 * 125
    assert a.x.min() == 0.0 and a.x.max() == 1.0
    assert abs(a.cell_volume.sum() - 1.0) <= 1e-13
    # jitter within +-10 % of the spacing
    grid = np.round(a.x * 5) / 5
    assert np.max(np.abs(a.x - grid)) <= 0.1 / 5 + 1e-15


@pytest.mark.parametrize("order", ["rcm", "morton"])
def test_renumbering_is_a_permutation(order):
    a = meshes.kuhn_tets(5, order=None)
    b = meshes.kuhn_tets(5, order=order)
    assert a.n_cells == b.n_cells and a.n_faces == b.n_faces and a.n_interior == b.n_interior
    ka = np.sort(np.round(a.centroids(), 12), axis=0)
    kb = np.sort(np.round(b.centroids(), 12), axis=0)
    np.testing.assert_array_equal(ka, kb)
    np.testing.assert_allclose(np.sort(a.cell_volume), np.sort(b.cell_volume), rtol=1e-14)


@pytest.mark.parametrize("name", list(FIXTURES))
def test_face2_stencil_contents(name):
    """codegen/cpp/geometry.cpp:432-617: the face neighbours and their face neighbours (not the
    cell), ascending, then the boundary faces of both, ascending, as -(f + 1)."""
    m = fixture(name)
    off, cf, own, nei, ni = m.cell_face_offsets, m.cell_faces, m.owner, m.neighbour, m.n_interior

    def faces(c):
        return cf[off[c] : off[c + 1]]

    def across(c, f):
        return nei[f] if own[f] == c else own[f]

    for c in range(m.n_cells):
        cells, bfaces = set(), set()
        for f in faces(c):
            if f >= ni:
                bfaces.add(f)
                continue
            nb = across(c, f)
            cells.add(nb)
            for f2 in faces(nb):
                if f2 >= ni:
                    bfaces.add(f2)
                else:
                    cells.add(across(nb, f2))
        cells.discard(c)
        expect = sorted(cells) + [-f - 1 for f in sorted(bfaces)]
        got = m.stencil[m.stencil_offsets[c] : m.stencil_offsets[c + 1]].tolist()
        assert got == expect


@pytest.mark.parametrize("name", list(FIXTURES))
def test_lsq_weights_invert_the_moment_matrix(name):
    """w_k = M^-1 d_k / |d_k|^2 with M = sum d d^T / |d|^2 over the used entries: sum_k w_k d_k^T is
    the identity wherever the fit is determined (every boundary face a `state` face)."""
    m = fixture(name)
    meshes.set_boundary_kinds(m, np.zeros(m.n_boundary, dtype=np.int32))
    xc = m.centroids()
    xf = m.face_centroids()
    w = m.stencil_weights.reshape(3, -1)
    for c in range(m.n_cells):
        k0, k1 = m.stencil_offsets[c], m.stencil_offsets[c + 1]
        e = m.stencil[k0:k1]
        pos = np.where(e[:, None] >= 0, xc[np.maximum(e, 0)], xf[np.maximum(-e - 1, 0)])
        d = pos - xc[c]
        ident = w[:, k0:k1] @ d
        np.testing.assert_allclose(ident, np.eye(3), atol=1e-10)


def test_boundary_faces_join_the_fit_only_as_state_faces():
    """physics/src/euler.cpp:340-416: slip faces give no weight while the interior entries
    determine the fit; the cells that cannot keep every boundary face."""
    m = fixture("tet-gmsh")
    meshes.set_boundary_kinds(m, np.full(m.n_boundary, meshes.SLIP, dtype=np.int32))
    w = m.stencil_weights.reshape(3, -1)
    bnd = m.stencil < 0
    kept = m.kept_boundaries
    counts = np.diff(m.stencil_offsets)
    cell_of = np.repeat(np.arange(m.n_cells), counts)
    # in the cells that did not need them, boundary entries weigh exactly zero
    weighted = np.unique(cell_of[bnd & np.any(w != 0, axis=0)])
    assert weighted.size == kept
    meshes.set_boundary_kinds(m, np.zeros(m.n_boundary, dtype=np.int32))
    assert m.kept_boundaries == 0
    assert np.any(m.stencil_weights.reshape(3, -1)[:, bnd] != 0)
