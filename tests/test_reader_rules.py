"""The mesh readers' remaining rules from the reference's readers:

* polyMesh boundary (io/src/polymesh.cpp:486-509): each patch dictionary read as a unit (any key
  order), each patch starting where the previous one ends, together covering every boundary face;
  processor patches (a decomposed mesh) refused as unsupported (E0133).
* polyMesh binary points in 32-bit scalars, widened to double (polymesh.cpp:226-231), and uniform
  label lists N{v} for owner and neighbour (polymesh.cpp:169-178).
* Gmsh periodic links: both surfaces in a physical group (io/src/gmsh.cpp:178-182); the periodic
  face's neighbour seen a period away on the far side of the face.
"""

from __future__ import annotations

import shutil
import struct
from pathlib import Path

import numpy as np
import pytest

from autan_amd import mesh as meshes

GOLDEN = Path(__file__).resolve().parent / "golden" / "meshes"


def copy(tmp_path: Path, name: str) -> Path:
    d = tmp_path / name
    shutil.copytree(GOLDEN / name, d)
    return d


def same_mesh(a, b) -> None:
    for k in a.arrays:
        assert np.array_equal(np.asarray(a.arrays[k]).view(np.uint8), np.asarray(b.arrays[k]).view(np.uint8)), k


@pytest.mark.parametrize("edit,code", [
    (("nFaces          9;\n        startFace       54;", "nFaces          18;\n        startFace       54;"), "E0132"),
    (("startFace       63;", "startFace       64;"), "E0132"),  # a gap in the boundary range
    (("nFaces          9;\n        startFace       99;", "nFaces          8;\n        startFace       99;"), "E0132"),
    (("type            wall;", "type            processor;"), "E0133"),
    (("type            patch;\n        physicalType    patch;\n        nFaces          9;\n        startFace       54;",
      "type            processorCyclic;\n        physicalType    patch;\n        nFaces          9;\n"
      "        startFace       54;"), "E0133"),
    (("type            wall;\n", ""), "E0132"),  # a patch without its type
    (("6 // entry0", "7 // entry0"), "E0132"),  # more patches announced than written
])
def test_boundary_patches_are_checked_as_the_reference_checks_them(tmp_path, edit, code):
    d = copy(tmp_path, "hex")
    text = (d / "boundary").read_text()
    assert edit[0] in text, edit[0]
    (d / "boundary").write_text(text.replace(edit[0], edit[1], 1))
    with pytest.raises(RuntimeError, match=code):
        meshes.read_polymesh(d)


def test_patch_keys_in_any_order_and_extra_entries(tmp_path):
    d = copy(tmp_path, "hex")
    text = (d / "boundary").read_text()
    text = text.replace("        nFaces          9;\n        startFace       72;",
                        "        startFace       72;\n        inGroups        List<word> 1(wall);\n"
                        "        sub { a 1; b (1 2); }\n        nFaces          9;")
    (d / "boundary").write_text(text)
    same_mesh(meshes.read_polymesh(d), meshes.read_polymesh(GOLDEN / "hex"))


def test_binary_points_in_32_bit_scalars_are_widened(tmp_path):
    """The binary hybrid mesh with its points rewritten as 32-bit floats: the mesh read equals the
    one read from the same points rounded to float and written as doubles."""
    src = (GOLDEN / "hybrid-binary" / "points").read_bytes()
    head_end = src.index(b"(", src.index(b"}")) + 1
    count = int(src[src.index(b"}") + 1 : head_end - 1].split()[-1])
    xyz = np.frombuffer(src[head_end : head_end + 24 * count], dtype="<f8")
    tail = src[head_end + 24 * count :]
    head = src[:head_end]
    as32 = xyz.astype("<f4")
    d32 = copy(tmp_path, "hybrid-binary")
    (d32 / "points").write_bytes(head.replace(b"scalar=64", b"scalar=32") + as32.tobytes() + tail)
    d64 = tmp_path / "rounded"
    shutil.copytree(GOLDEN / "hybrid-binary", d64)
    (d64 / "points").write_bytes(head + as32.astype("<f8").tobytes() + tail)
    same_mesh(meshes.read_polymesh(d32), meshes.read_polymesh(d64))


HEX_POINTS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
HEX_FACES = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (0, 4, 7, 3), (1, 2, 6, 5)]


def foam(cls: str, obj: str, body: str) -> str:
    return (f"FoamFile\n{{\n    version 2.0;\n    format ascii;\n    class {cls};\n    object {obj};\n}}\n\n"
            f"{body}\n")


def one_hex(d: Path, owner: str, neighbour: str) -> Path:
    d.mkdir()
    pts = "\n".join(f"({x} {y} {z})" for x, y, z in HEX_POINTS)
    (d / "points").write_text(foam("vectorField", "points", f"8\n(\n{pts}\n)"))
    fcs = "\n".join(f"4({' '.join(map(str, f))})" for f in HEX_FACES)
    (d / "faces").write_text(foam("faceList", "faces", f"6\n(\n{fcs}\n)"))
    (d / "owner").write_text(foam("labelList", "owner", owner))
    (d / "neighbour").write_text(foam("labelList", "neighbour", neighbour))
    (d / "boundary").write_text(foam("polyBoundaryMesh", "boundary",
                                     "1\n(\n    walls\n    {\n        type wall;\n        nFaces 6;\n"
                                     "        startFace 0;\n    }\n)"))
    return d


def test_uniform_label_lists(tmp_path):
    """A one-cell mesh's owner written 6{0} reads as the same mesh as 6(0 0 0 0 0 0); a uniform list
    longer than the faces is refused."""
    a = meshes.read_polymesh(one_hex(tmp_path / "a", "6{0}", "0()"))
    b = meshes.read_polymesh(one_hex(tmp_path / "b", "6(0 0 0 0 0 0)", "0()"))
    same_mesh(a, b)
    assert a.n_cells == 1 and abs(a.cell_volume.sum() - 1.0) < 1e-15
    with pytest.raises(RuntimeError, match="E0132"):
        meshes.read_polymesh(one_hex(tmp_path / "c", "7{0}", "0()"))


PERIODIC = GOLDEN / "gmsh" / "periodic-box.ascii.msh"


def test_a_periodic_surface_in_no_physical_group_is_refused(tmp_path):
    text = PERIODIC.read_text()
    # surface entity 2 (xmax, the slave of the link "2 2 1") loses its physical group
    old = "1.0000001 1.0000001 1.0000001 1 2 4 -5 8 7 -6"
    assert old in text and "\n2 2 1\n" in text
    path = tmp_path / "p.msh"
    path.write_text(text.replace(old, "1.0000001 1.0000001 1.0000001 0 4 -5 8 7 -6", 1))
    with pytest.raises(RuntimeError, match="E0132: gmsh: periodic surface 2 is in no physical group"):
        meshes.read_gmsh(path)


@pytest.mark.parametrize("encoding", ["ascii", "binary"])
def test_periodic_neighbours_lie_beyond_the_face(encoding):
    """Every interior face, periodic ones included, has its owner's centroid behind it and its
    neighbour's image in front of it: n . r_left > 0 and n . r_right < 0, with |r_right| of the
    order of a cell (a shift of the wrong sign would put the image a period away, |r| ~ 1)."""
    m = meshes.read_gmsh(GOLDEN / "gmsh" / f"periodic-box.{encoding}.msh")
    ni = m.n_interior
    n = m.arrays["face_normal"].reshape(3, -1)[:, :ni]
    rl = m.arrays["r_left"].reshape(3, -1)[:, :ni]
    rr = m.arrays["r_right"].reshape(3, -1)[:, :ni]
    assert np.all((n * rl).sum(axis=0) > 0)
    assert np.all((n * rr).sum(axis=0) < 0)
    h = m.cell_volume.max() ** (1.0 / 3.0)
    assert np.linalg.norm(rr, axis=0).max() < 3 * h < 0.5
    moved = np.flatnonzero(np.abs(m.arrays["shift"].reshape(3, -1)[:, :ni]).sum(axis=0))
    assert moved.size > 0
