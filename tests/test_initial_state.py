"""autan.initial-state/1 files (autan_amd/initial_state.py) with the cases of the reference's own
tests: the SDK's writer and its refusals (py/tests/test_initial_state.py), the runtime's checks
and pairing (tests/runtime/test_runtime.cpp:2051-2236: E0508 to E0511, rows moved by 0.999 of the
tolerance, one row within the tolerance of two cells), and the start state set_cell builds."""

from __future__ import annotations

import hashlib
import json

import numpy as np
import pytest

from autan_amd import initial_state as ist
from autan_amd import mesh as meshes

GAMMA = 1.4


@pytest.fixture(scope="module")
def tets():
    m = meshes.kuhn_tets(4, 0.1, 0, order="rcm")
    points = np.column_stack([m.x, m.y, m.z])
    return m, points


def field(n: int) -> np.ndarray:
    """The runtime test's rows: density 1 + 1e-3 c, velocity (c, -c, 0.5), pressure 1e5 + c."""
    c = np.arange(n, dtype=np.float64)
    return np.column_stack([1.0 + 1e-3 * c, c, -c, np.full(n, 0.5), 1e5 + c])


def test_the_written_file_is_the_documented_json(tmp_path):
    """The SDK's own case: the document the contract shows, in the bytes Python's json module
    writes for it, and its SHA-256 returned."""
    path = tmp_path / "s.json"
    centroids = [[0.1, 0.2, 0.3], [-1.0, 2.0, 3.5]]
    values = [[1.2, 30.0, -4.0, 0.5, 1e5], [0.9, -1.0, 2.0, 0.0, 9e4]]
    digest = ist.write(path, oracle="vortex", mesh_sha256="ab" * 32, centroids=centroids, values=values)
    want = {
        "schema": "autan.initial-state/1",
        "oracle": "vortex",
        "mesh_sha256": "ab" * 32,
        "variables": ["density", "velocity-x", "velocity-y", "velocity-z", "pressure"],
        "cells": 2,
        "centroids": centroids,
        "values": values,
    }
    data = path.read_bytes()
    assert json.loads(data) == want and list(json.loads(data)) == list(want)
    assert data == json.dumps(want).encode()
    assert digest == hashlib.sha256(data).hexdigest()
    assert data.startswith(b'{"schema": "autan.initial-state/1", "oracle": "vortex", "mesh_sha256": "abab')


@pytest.mark.parametrize("values", [
    [[1.2, 30.0, 0.0, 0.0, float("nan")]] * 4,
    [[-1.2, 30.0, 0.0, 0.0, 1e5]] * 4,
    [[1.2, 30.0, 0.0, 1e5]] * 4,
    [[1.2, 30.0, 0.0, 0.0, 1e5]] * 3,
])
def test_write_refuses_what_the_contract_refuses(tmp_path, values):
    centroids = [[0.1 * i, 0.0, 0.0] for i in range(4)]
    with pytest.raises(ValueError):
        ist.write(tmp_path / "s.json", oracle="o", mesh_sha256="0" * 64, centroids=centroids, values=values)


def test_write_refuses_transposed_rows(tmp_path):
    centroids = np.zeros((3, 4))  # 3 x n instead of n x 3
    with pytest.raises(ist.InitialStateError, match="E0508"):
        ist.write(tmp_path / "s.json", oracle="o", mesh_sha256="0" * 64, centroids=centroids,
                  values=[[1.2, 30.0, 0.0, 0.0, 1e5]] * 4)


def _doc(n=4) -> dict:
    return {"schema": ist.SCHEMA, "oracle": "o", "mesh_sha256": "0" * 64, "variables": list(ist.VARIABLES),
            "cells": n, "centroids": [[0.1 * i, 0.0, 0.0] for i in range(n)],
            "values": [[1.2, 30.0, 0.0, 0.0, 100000.0] for _ in range(n)]}


def _text(**changes) -> bytes:
    return json.dumps(_doc() | changes).encode()


def _padded(chars: int) -> bytes:
    """The first pressure spelt in `chars` characters, the same value."""
    return _text().replace(b"100000.0", b"100000." + b"0" * (chars - 7), 1)


@pytest.mark.parametrize("data,ok", [
    (b"{}", False),
    (b'{"schema": "autan.initial-state/1"}', False),
    (_text().replace(b"100000.0", b"NaN", 1), False),
    (_text().replace(b'"cells":', b'"cells": 4, "cells":', 1), False),
    (_text(oracle="é" * 2048), True),
    (_text(oracle="é" * 2048 + "x"), False),
    (_text(oracle="o" * 4096), True),
    (_text(oracle="o" * 4097), False),
    (_padded(64), True),
    (_padded(65), False),
    (b"[" * 100_000 + b"]" * 100_000, False),
    (_text(note="unknown"), False),
    (_text(mesh_sha256="not-a-sha256"), False),
    (_text(variables=["density", "velocity-x", "velocity-y", "velocity-z", "temperature"]), False),
    (_text(values=[[0.0, 30.0, 0.0, 0.0, 1e5]] + [[1.2, 30.0, 0.0, 0.0, 1e5]] * 3), False),
    (_text(centroids=[[{"values": []}]] + [[0.1, 0.0, 0.0]] * 3), False),
    (_text(cells=4.0), False),
    (_text(cells=True), False),
])
def test_the_file_is_checked(data, ok):
    if ok:
        st = ist.parse(data)
        assert st.values.shape == (4, 5) and st.sha256 == hashlib.sha256(data).hexdigest()
    else:
        with pytest.raises(ist.InitialStateError) as e:
            ist.parse(data)
        assert e.value.code == "E0508"


def test_the_files_hash_and_mesh(tmp_path):
    data = _text()
    with pytest.raises(ist.InitialStateError, match="E0508"):
        ist.parse(data, sha256="0" * 64)
    assert ist.parse(data, sha256=hashlib.sha256(data).hexdigest()).oracle == "o"
    with pytest.raises(ist.InitialStateError) as e:
        ist.parse(data, mesh_sha256="a" * 64)
    assert e.value.code == "E0509"
    with pytest.raises(ist.InitialStateError) as e:
        ist.read(tmp_path / "missing.json")
    assert e.value.code == "E0508"


def test_rows_in_any_order_give_each_cell_its_state(tmp_path, tets):
    """Rows shuffled and each moved by 0.999 of the tolerance along x, y or z in turn: still their
    cells' rows; the start state is set_cell's of each cell's row."""
    m, points = tets
    x = m.centroids()
    n = x.shape[0]
    vals = field(n)
    perm = np.random.default_rng(3).permutation(n)
    rows = x[perm].copy()
    tol = ist.TOLERANCE * ist.extent(points)
    for r in range(n):
        rows[r, r % 3] += 0.999 * tol
    digest = ist.write(tmp_path / "s.json", oracle="field", mesh_sha256="c" * 64, centroids=rows, values=vals[perm])
    st = ist.read(tmp_path / "s.json", sha256=digest, mesh_sha256="c" * 64)
    match = ist.pair(st.centroids, x, ist.extent(points))
    assert np.array_equal(perm[match], np.arange(n))
    u = ist.start_state(st, x, points, GAMMA).reshape(5, n)
    c = np.arange(n, dtype=np.float64)
    rho = 1.0 + 1e-3 * c
    assert np.array_equal(u[0], rho) and np.array_equal(u[1], rho * c) and np.array_equal(u[2], rho * -c)
    assert np.array_equal(u[3], rho * 0.5)
    assert np.array_equal(u[4], (1e5 + c) / (GAMMA - 1.0) + 0.5 * rho * (c * c + c * c + 0.25))


def test_a_row_short_or_one_row_is_e0510(tets):
    m, points = tets
    x = m.centroids()
    for rows in (x[:-1], x[:1]):
        with pytest.raises(ist.InitialStateError) as e:
            ist.pair(rows, x, ist.extent(points))
        assert e.value.code == "E0510"


def test_a_far_row_or_two_rows_on_one_cell_is_e0511(tets):
    m, points = tets
    x = m.centroids()
    far = x.copy()
    far[3, 0] += 1e-3
    twice = x.copy()
    twice[3] = twice[4]
    for rows in (far, twice):
        with pytest.raises(ist.InitialStateError) as e:
            ist.pair(rows, x, ist.extent(points))
        assert e.value.code == "E0511"


def test_one_row_within_the_tolerance_of_two_cells_is_e0511():
    """Cells 1.5e-9 apart in a box of extent 1 (tolerance 1e-9): each within the tolerance of its
    own row only, until cell 0's row moves half-way to the other cell and the other's row away."""
    cells = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.5e-9], [1.0, 1.0, 1.0]])
    points = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    assert np.array_equal(ist.pair(cells.copy(), cells, ist.extent(points)), [0, 1, 2])
    rows = cells.copy()
    rows[0] = 0.5 * (cells[0] + cells[1])
    rows[1] = [-1.0, -1.0, -1.0]
    with pytest.raises(ist.InitialStateError) as e:
        ist.pair(rows, cells, ist.extent(points))
    assert e.value.code == "E0511" and "match one row" in str(e.value)


def test_two_cells_within_the_tolerance_cannot_be_paired():
    cells = np.array([[0.0, 0.0, 0.0], [0.0, 0.5e-9, 0.0], [1.0, 1.0, 1.0]])
    with pytest.raises(ist.InitialStateError, match="cells 0 and 1 lie within"):
        ist.pair(cells.copy(), cells, 1.0)
