"""Gmsh MSH 4.1 meshes, ASCII and binary (mxmesh_read_gmsh, io/src/gmsh.cpp), against the same meshes
read from polyMesh: the reference's test meshes in both formats (tests/data/mesh/<name>/, copied under
tests/golden/meshes/), which its own test reads to the same mesh (tests/io/test_readers.cpp:30-74);
a hand-written file from the format's specification (test_readers.cpp:207-229); the refusals of
test_readers.cpp:230-270."""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd.stepper import freestream_scheme

GOLDEN = Path(__file__).resolve().parent / "golden" / "meshes"
NAMES = ("hex", "hybrid", "tet-gmsh")


def polymesh_zones(name: str) -> list[str]:
    """Patch names of a polyMesh boundary file, in file order (the reader's zone order)."""
    text = (GOLDEN / name / "boundary").read_text()
    body = text[text.index("(", text.index("}") + 1) + 1 :]
    return re.findall(r"^\s*([A-Za-z_][\w.-]*)\s*\n\s*\{", body, flags=re.M)


def gmsh_zones(path: Path) -> list[str]:
    """Names of the 2-D physical groups by increasing tag (the reader's zone order)."""
    text = path.read_text()
    body = text[text.index("$PhysicalNames") : text.index("$EndPhysicalNames")].splitlines()[2:]
    groups = [(int(t), n.strip('"')) for d, t, n in (line.split(None, 2) for line in body) if d == "2"]
    return [n for _, n in sorted(groups)]


def by_centroid(points: np.ndarray) -> np.ndarray:
    r = np.round(points, 9)
    return np.lexsort((r[:, 2], r[:, 1], r[:, 0]))


def read_both(name: str, encoding: str = "ascii"):
    return meshes.read_gmsh(GOLDEN / "gmsh" / f"{name}.{encoding}.msh"), meshes.read_polymesh(GOLDEN / name)


@pytest.mark.parametrize("encoding", ["ascii", "binary"])
@pytest.mark.parametrize("name", NAMES)
def test_every_format_reads_to_the_same_mesh(name, encoding):
    g, p = read_both(name, encoding)
    assert (g.n_cells, g.n_faces, g.n_interior) == (p.n_cells, p.n_faces, p.n_interior)
    og, op = by_centroid(g.centroids()), by_centroid(p.centroids())
    np.testing.assert_allclose(g.centroids()[og], p.centroids()[op], rtol=0, atol=1e-13)
    np.testing.assert_allclose(g.cell_volume[og], p.cell_volume[op], rtol=1e-12)
    zg = gmsh_zones(GOLDEN / "gmsh" / f"{name}.ascii.msh")
    zp = polymesh_zones(name)
    assert sorted(zg) == sorted(zp) and len(zp) == 6
    xg, xp = g.face_centroids()[g.n_interior :], p.face_centroids()[p.n_interior :]
    ag = g.face_area[g.n_interior :]
    ap = p.face_area[p.n_interior :]
    for zone in zp:
        sg = g.face_zone == zg.index(zone)
        sp = p.face_zone == zp.index(zone)
        assert sg.sum() == sp.sum() > 0, zone
        fg, fp = by_centroid(xg[sg]), by_centroid(xp[sp])
        np.testing.assert_allclose(xg[sg][fg], xp[sp][fp], rtol=0, atol=1e-13)
        np.testing.assert_allclose(ag[sg][fg], ap[sp][fp], rtol=1e-12)


@pytest.mark.parametrize("name", NAMES)
def test_binary_and_ascii_files_hold_the_same_mesh(name):
    """The same topology and zones; the points equal to the ASCII file's printed digits (Gmsh
    writes coordinates there with fewer digits than a double round-trips)."""
    a = meshes.read_gmsh(GOLDEN / "gmsh" / f"{name}.ascii.msh")
    b = meshes.read_gmsh(GOLDEN / "gmsh" / f"{name}.binary.msh")
    for k in ("owner", "neighbour", "face_node_offsets", "face_nodes", "face_zone", "cell_face_offsets"):
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(a.arrays[k], b.arrays[k], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", NAMES)
def test_the_residual_is_the_same_on_both_readings(name):
    """The oracle's second-order residual of one smooth field on both readings, cell by cell: the
    same geometry, stencils and weights up to summation order."""
    from oracle.oracle import Oracle

    out = []
    for m in read_both(name):
        kinds = np.zeros(m.n_boundary, dtype=np.int32)
        meshes.set_boundary_kinds(m, kinds)
        sc = freestream_scheme(1.0, (0.4, 0.1, -0.05), 1 / 1.4, flux="roe", second_order=True, limiter=True)
        bc = cases.state_faces_data(m, lambda x: cases._prim(cases.smooth_state(x)))
        u = cases.smooth_state(m.centroids())
        r = Oracle(m, sc, "fp64", kinds.astype(np.float64), bc).residual(u.reshape(-1)).reshape(5, -1)
        out.append(r[:, by_centroid(m.centroids())])
    scale = np.abs(out[1]).max(axis=1, keepdims=True)
    assert np.all(scale > 0)
    np.testing.assert_allclose(out[0] / scale, out[1] / scale, rtol=0, atol=1e-12)


# A unit tetrahedron written by hand from the MSH 4.1 specification: its base (z = 0) in physical
# surface 7 "base", the other three faces in 8 "sides", the volume in 9 "solid-block".
TET = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
3
2 7 "base"
2 8 "sides"
3 9 "solid-block"
$EndPhysicalNames
$Entities
0 0 2 1
1 0 0 0 1 1 0 1 7 0
2 0 0 0 1 1 1 1 8 0
1 0 0 0 1 1 1 1 9 2 1 2
$EndEntities
$Nodes
1 4 1 4
3 1 0 4
1
2
3
4
0 0 0
1 0 0
0 1 0
0 0 1
$EndNodes
$Elements
3 5 1 6
2 1 2 1
1 1 3 2
2 2 2 3
2 1 2 4
3 2 3 4
4 1 4 3
3 1 4 1
6 NODES
$EndElements
"""


def write(tmp_path: Path, text: str) -> Path:
    p = tmp_path / "m.msh"
    p.write_text(text)
    return p


@pytest.mark.parametrize("nodes", ["1 2 3 4", "1 3 2 4"])
def test_a_hand_written_tetrahedron(tmp_path, nodes):
    """One cell of volume 1/6 in either node orientation, its base alone in zone 0 and the three
    other faces in zone 1."""
    m = meshes.read_gmsh(write(tmp_path, TET.replace("6 NODES", f"6 {nodes}")))
    assert (m.n_cells, m.n_faces, m.n_interior) == (1, 4, 0)
    assert abs(m.cell_volume[0] - 1 / 6) < 1e-15
    assert sorted(m.face_zone.tolist()) == [0, 1, 1, 1]
    base = m.face_centroids()[m.face_zone == 0]
    np.testing.assert_allclose(base, [[1 / 3, 1 / 3, 0.0]], atol=1e-15)
    normal = m.face_normal.reshape(3, -1)[:, m.face_zone == 0].ravel()
    np.testing.assert_allclose(normal, [0.0, 0.0, -1.0], atol=1e-15)  # out of the cell


@pytest.mark.parametrize("change,code", [
    (("4.1 0 8", "4.1 0 4"), "E0133"),
    (("4.1 0 8", "4.1 2 8"), "E0132"),
    (("4.1 0 8", "2.2 0 8"), "E0133"),
    (("3 1 4 1", "3 1 11 1"), "E0133"),
    (("$EndElements\n", "$EndElements\n$Periodic\n1\n2 1 2\n16 0 -1 0 0 1 0 0 0 0 0 1 0 0 0 0 1\n0\n$EndPeriodic\n"),
     "E0133"),
    (("2 0 0 0 1 1 1 1 8 0", "2 0 0 0 1 1 1 2 7 8 0"), "E0135"),
    (("$MeshFormat", "$Format"), "E0131"),
    (("6 NODES", "6 1 2 3 99"), "E0132"),
])
def test_refusals(tmp_path, change, code):
    text = TET.replace(*change).replace("6 NODES", "6 1 2 3 4")
    with pytest.raises(RuntimeError, match=code):
        meshes.read_gmsh(write(tmp_path, text))


def test_a_missing_file_is_e0130(tmp_path):
    with pytest.raises(RuntimeError, match="E0130"):
        meshes.read_gmsh(tmp_path / "none.msh")


@pytest.mark.gpu
@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_the_device_residual_on_a_gmsh_mesh_is_the_oracles(policy):
    """The hybrid mesh (tetrahedra, hexahedra, prisms, pyramids) read from Gmsh, every boundary a
    `state` face: the device residual equals the oracle's bit for bit."""
    from autan_amd.stepper import DeviceStepper
    from oracle.oracle import Oracle

    m = meshes.read_gmsh(GOLDEN / "gmsh" / "hybrid.ascii.msh", order="rcm")
    kinds = np.zeros(m.n_boundary, dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    sc = freestream_scheme(1.0, (0.4, 0.1, -0.05), 1 / 1.4, flux="hllc", second_order=True, limiter=True)
    bc = cases.state_faces_data(m, lambda x: cases._prim(cases.smooth_state(x)))
    u = cases.smooth_state(m.centroids()).reshape(-1)
    print(f"gmsh hybrid {policy}", flush=True)
    dev = DeviceStepper(m, sc, policy, kinds.astype(np.float64), bc)
    got = dev.residual(u)
    dev.close()
    want = Oracle(m, sc, policy, kinds.astype(np.float64), bc).residual(u)
    assert np.array_equal(np.asarray(got).reshape(-1).view(np.uint64), np.asarray(want).reshape(-1).view(np.uint64))


@pytest.mark.parametrize("encoding", ["ascii", "binary"])
def test_damaged_files_are_errors_not_crashes(tmp_path, encoding):
    """Truncations and single-byte changes of a test mesh: each either reads or raises the reader's
    error (the counts a file declares are checked against its size before anything is allocated,
    tests/io/test_readers.cpp:386-426)."""
    data = (GOLDEN / "gmsh" / f"hex.{encoding}.msh").read_bytes()
    rng = np.random.default_rng(12)
    path = tmp_path / "d.msh"
    outcomes = {"read": 0, "refused": 0}
    cases_ = [data[: int(k)] for k in rng.integers(0, len(data), 150)]
    for k in rng.integers(0, len(data), 150):
        b = bytearray(data)
        b[int(k)] = int(rng.integers(0, 256))
        cases_.append(bytes(b))
    for d in cases_:
        path.write_bytes(d)
        try:
            meshes.read_gmsh(path)
            outcomes["read"] += 1
        except RuntimeError:
            outcomes["refused"] += 1
    assert outcomes["refused"] > 0 and sum(outcomes.values()) == 300


@pytest.mark.parametrize("name", ["hybrid", "poly"])
def test_binary_polymesh_reads_as_the_ascii_one(name):
    """OpenFOAM's binary polyMesh (labels and scalars raw, the faces as a compact list) of the
    reference's test meshes: the ASCII directory's mesh, bit for bit (test_readers.cpp:42-43)."""
    a = meshes.read_polymesh(GOLDEN / name)
    b = meshes.read_polymesh(GOLDEN / f"{name}-binary")
    for k in a.arrays:
        assert np.array_equal(np.asarray(a.arrays[k]).view(np.uint8), np.asarray(b.arrays[k]).view(np.uint8)), k


@pytest.mark.parametrize("source", ["hex", "hybrid-binary"])
def test_damaged_polymesh_files_are_errors_not_crashes(tmp_path, source):
    """Truncations and single-byte changes of each file of a polyMesh: the reader refuses with its
    error or reads a mesh; every count and index is checked before it is used."""
    import shutil

    rng = np.random.default_rng(5)
    refused = 0
    for trial in range(150):
        d = tmp_path / f"m{trial}"
        shutil.copytree(GOLDEN / source, d)
        name = ("points", "faces", "owner", "neighbour", "boundary")[trial % 5]
        data = bytearray((d / name).read_bytes())
        if trial % 2:
            data = data[: int(rng.integers(0, len(data)))]
        else:
            data[int(rng.integers(0, len(data)))] = int(rng.integers(0, 256))
        (d / name).write_bytes(bytes(data))
        try:
            meshes.read_polymesh(d)
        except RuntimeError:
            refused += 1
    assert refused > 0



PERIODIC = GOLDEN / "gmsh" / "periodic-box.ascii.msh"


@pytest.mark.parametrize("encoding", ["ascii", "binary"])
def test_periodic_surface_links_become_periodic_faces(encoding):
    """The reference's periodic box (tetrahedra in the unit cube, xmax meshed as xmin moved by
    (1, 0, 0), ymax as ymin by (0, 1, 0); tests/io/test_readers.cpp:428-444): those faces become
    interior faces whose neighbour lies a period away, every cell stays closed, and only zmin and
    zmax stay boundary zones."""
    m = meshes.read_gmsh(GOLDEN / "gmsh" / f"periodic-box.{encoding}.msh")
    zones = gmsh_zones(PERIODIC)  # the binary file's names are the ASCII file's
    assert sorted(zones[z] for z in set(m.face_zone.tolist())) == ["zmax", "zmin"]
    shift = m.arrays["shift"].reshape(3, -1)
    moved = np.flatnonzero(np.abs(shift).sum(axis=0))
    assert moved.size > 0
    assert {tuple(np.abs(shift[:, f])) for f in moved} == {(1.0, 0.0, 0.0), (0.0, 1.0, 0.0)}
    av = m.arrays["face_area_vector"].reshape(3, -1)
    closed = np.zeros((m.n_cells, 3))
    np.add.at(closed, m.owner, av.T)
    np.add.at(closed, m.neighbour, -av[:, : m.n_interior].T)
    assert np.abs(closed).max() < 1e-15
    assert abs(m.cell_volume.sum() - 1.0) < 1e-13
    # Every interior face, periodic ones included, lies between its cells as the owner sees them:
    # the owner behind it (n . r_left > 0), the neighbour's image (its centroid plus the face's
    # shift) in front (n . r_right < 0), both within about a cell; a shift of the wrong sign puts
    # the image two periods off.
    ni = m.n_interior
    xf = m.face_centroids()[:ni]
    xc = m.centroids()
    n = av[:, :ni].T
    r_left = xf - xc[m.owner[:ni]]
    r_right = xf - (xc[m.neighbour[:ni]] + shift[:, :ni].T)
    assert np.all(np.einsum("ij,ij->i", n, r_left) > 0)
    assert np.all(np.einsum("ij,ij->i", n, r_right) < 0)
    h = np.cbrt(m.cell_volume.max())
    assert np.linalg.norm(r_right[moved[moved < ni]], axis=1).max() < 2 * h


def test_a_uniform_flow_is_steady_on_the_periodic_box():
    """Every face, periodic ones included, oriented and placed right: a uniform flow with its own
    state on zmin and zmax has a zero residual (to round-off) in every cell."""
    from oracle.oracle import Oracle

    m = meshes.read_gmsh(PERIODIC)
    kinds = np.zeros(m.n_boundary, dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    w = np.array([1.0, 0.3, 0.2, 0.1, 1.0 / 1.4])
    sc = freestream_scheme(w[0], w[1:4], w[4], flux="roe", second_order=True, limiter=True)
    bc = np.repeat(w[:, None], m.n_boundary, axis=1)
    u = np.repeat(np.asarray(cases.conserved(*w))[:, None], m.n_cells, axis=1)
    r = Oracle(m, sc, "fp64", kinds.astype(np.float64), bc).residual(u.reshape(-1)).reshape(5, -1)
    assert np.abs(r).max() < 1e-14


@pytest.mark.parametrize("edit,code", [
    (("2 2 1\n16 1 0 0 1 0 1 0 0", "2 2 1\n16 0 -1 0 1 1 0 0 0"), "E0133"),  # a rotation
    (("2 2 1\n16 1 0 0 1 0", "2 2 1\n16 1 0 0 1.01 0"), "E0132"),  # node pairs off the translation
    (("58\n5 1\n6 2\n", "58\n5 2\n6 1\n"), "E0132"),  # a node pair swapped
    (("2 2 1\n16 1 0 0 1", "4 2 1\n16 1 0 0 1"), "E0132"),  # a dimension the format lacks
    (("\n1 3 1\n", "\n1 0 1\n"), "E0132"),  # a link to entity tag 0
    (("2 2 1\n16 1 0 0 1 0 1 0 0 0 0 1 0 0 0 0 1\n", "2 2 1\n0\n"), "E0133"),  # no affine map
    (("$Periodic\n15\n", "$Periodic\n16\n2 2 1\n16 1 0 0 2 0 1 0 0 0 0 1 0 0 0 0 1\n0\n"), "E0132"),
    (("2 2 1\n16 1 0 0 1 0 1 0 0", "2 2 1\n16 1 0 0 1 0 1 0 1e-09"), None),  # off by the tolerance
])
def test_periodic_links_are_checked(tmp_path, edit, code):
    """The reference's edits of the periodic box's links (tests/io/test_readers.cpp:446-481)."""
    text = PERIODIC.read_text()
    assert edit[0] in text
    path = tmp_path / "p.msh"
    path.write_text(text.replace(edit[0], edit[1], 1))
    if code is None:
        assert meshes.read_gmsh(path).n_cells > 0
    else:
        with pytest.raises(RuntimeError, match=code):
            meshes.read_gmsh(path)
