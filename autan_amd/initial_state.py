"""autan.initial-state/1 files (docs/contracts/initial-state.md): the state of every fluid cell of
one mesh, cells identified by their centroid. The warm start's interchange with the reference's
SDK and runtime: `write` gives the bytes py/autan/initial_state.py:92-115 writes for the same rows
(so the same SHA-256), `read` makes the runtime's checks (runtime/src/run.cpp:999-1066, E0508 and
E0509), `pair` its pairing of rows with cells (run.cpp:1067-1178, E0510 and E0511), and
`start_state` the conserved state EulerPhysics::set_cell builds from the paired rows
(physics/src/euler_physics.cpp:795-816)."""

from __future__ import annotations

import hashlib
import json
import math
import re
from dataclasses import dataclass
from pathlib import Path

import numpy as np

SCHEMA = "autan.initial-state/1"
VARIABLES = ("density", "velocity-x", "velocity-y", "velocity-z", "pressure")
KEYS = ("schema", "oracle", "mesh_sha256", "variables", "cells", "centroids", "values")
TEXT_BYTES = 4096  # a string or member name, in UTF-8
NUMBER_CHARS = 64  # a number's text
TOLERANCE = 1e-9  # times the largest extent of the mesh's bounding box
_SHA256 = re.compile(r"[0-9a-f]{64}")


class InitialStateError(ValueError):
    """A file or rows the contract refuses: `code` is the runtime's diagnostic (E0508 not a valid
    document, E0509 another mesh, E0510 not one row per fluid cell, E0511 no one-to-one pairing)."""

    def __init__(self, code: str, message: str) -> None:
        super().__init__(f"{code}: {message}")
        self.code = code


@dataclass
class InitialState:
    oracle: str
    mesh_sha256: str
    centroids: np.ndarray  # rows x 3
    values: np.ndarray  # rows x 5: density, velocity x y z, pressure
    sha256: str  # of the file's bytes


def _text(s) -> bool:
    return isinstance(s, str) and len(s.encode("utf-8", "surrogatepass")) <= TEXT_BYTES


def _finite(x) -> bool:
    return isinstance(x, (int, float)) and not isinstance(x, bool) and math.isfinite(x)


def _problem(doc) -> str | None:
    """Why `doc` is not a valid document, or None."""
    if not isinstance(doc, dict):
        return "not a JSON object"
    for k in doc:
        if k not in KEYS:
            return f"unknown key '{k}'"
    if doc.get("schema") != SCHEMA:
        return "schema"
    if not _text(doc.get("oracle")):
        return "oracle"
    sha = doc.get("mesh_sha256")
    if not (isinstance(sha, str) and _SHA256.fullmatch(sha)):
        return "mesh_sha256"
    if doc.get("variables") != list(VARIABLES):
        return "variables"
    cells = doc.get("cells")
    if not isinstance(cells, int) or isinstance(cells, bool) or cells < 1:
        return "cells"
    for key, width in (("centroids", 3), ("values", 5)):
        rows = doc.get(key)
        if not isinstance(rows, list) or len(rows) != cells:
            return key
        for row in rows:
            if not isinstance(row, list) or len(row) != width or not all(_finite(x) for x in row):
                return key
    for row in doc["values"]:
        if not (row[0] > 0.0 and row[4] > 0.0):
            return "values: density and pressure must be > 0"
    return None


def document_bytes(*, oracle: str, mesh_sha256: str, centroids, values) -> bytes:
    """The file's bytes for these rows: the contract's keys in its order, numbers as Python writes
    a float, ", " and ": " separators, ASCII with escapes (py/autan/initial_state.py:104-114)."""
    # Row by row, as the SDK converts them: a row of the wrong length is refused, not reshaped.
    c = [[float(x) for x in row] for row in centroids]
    v = [[float(x) for x in row] for row in values]
    doc = {"schema": SCHEMA, "oracle": oracle, "mesh_sha256": mesh_sha256, "variables": list(VARIABLES),
           "cells": len(c), "centroids": c, "values": v}
    why = _problem(doc)
    if why is not None:
        raise InitialStateError("E0508", f"not an {SCHEMA} document: {why}")
    return json.dumps(doc).encode()


def write(path, *, oracle: str, mesh_sha256: str, centroids, values) -> str:
    """Writes the rows (one centroid and one state per fluid cell, in any order) and returns the
    file's SHA-256, which a case gives with the path."""
    data = document_bytes(oracle=oracle, mesh_sha256=mesh_sha256, centroids=centroids, values=values)
    Path(path).write_bytes(data)
    return hashlib.sha256(data).hexdigest()


def _number(convert):
    def checked(text: str):
        if len(text) > NUMBER_CHARS:
            raise ValueError(f"a number of {len(text)} characters, more than {NUMBER_CHARS}")
        return convert(text)

    return checked


def _members(pairs):
    names = [k for k, _ in pairs]
    for k in names:
        if not _text(k):
            raise ValueError(f"a member name of more than {TEXT_BYTES} bytes")
        if names.count(k) > 1:
            raise ValueError(f"member '{k}' appears twice")
    return dict(pairs)


def _constant(name: str):
    raise ValueError(f"{name} is not JSON")


def parse(data: bytes, sha256: str | None = None, mesh_sha256: str | None = None) -> InitialState:
    """The document in `data`, checked as the runtime checks a case's file: its SHA-256 when given,
    a valid document (E0508), written for the mesh whose hash is `mesh_sha256` when given (E0509)."""
    found = hashlib.sha256(data).hexdigest()
    if sha256 is not None and found != sha256:
        raise InitialStateError("E0508", "the initial-state file's SHA-256 differs from the case's")
    try:
        doc = json.loads(data, object_pairs_hook=_members, parse_float=_number(float), parse_int=_number(int),
                         parse_constant=_constant)
    except RecursionError:
        raise InitialStateError("E0508", f"not an {SCHEMA} document: nested too deeply") from None
    except ValueError as e:
        raise InitialStateError("E0508", f"not an {SCHEMA} document: {e}") from None
    why = _problem(doc)
    if why is not None:
        raise InitialStateError("E0508", f"not an {SCHEMA} document: {why}")
    if mesh_sha256 is not None and doc["mesh_sha256"] != mesh_sha256:
        raise InitialStateError("E0509", "the initial state was written for another mesh")
    return InitialState(doc["oracle"], doc["mesh_sha256"], np.array(doc["centroids"], dtype=np.float64),
                        np.array(doc["values"], dtype=np.float64), found)


def read(path, sha256: str | None = None, mesh_sha256: str | None = None) -> InitialState:
    p = Path(path)
    if not p.is_file():
        raise InitialStateError("E0508", f"'{p}' is not a file")
    return parse(p.read_bytes(), sha256, mesh_sha256)


def extent(points: np.ndarray) -> float:
    """The largest extent of the bounding box of the mesh's points (n x 3)."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    return float(np.max(p.max(axis=0) - p.min(axis=0)))


def _near(a: np.ndarray, b: np.ndarray, tol: float) -> np.ndarray:
    d = a - b
    return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) <= tol


def pair(rows: np.ndarray, cells: np.ndarray, box_extent: float) -> np.ndarray:
    """The row of each cell: every cell within the tolerance (1e-9 times the box's largest extent)
    of exactly one row, every row of exactly one cell, no two cells within the tolerance of each
    other. A distance is the square root of the sum of squares (the runtime takes std::hypot: the
    two agree but for a point within an ulp of the tolerance)."""
    from scipy.spatial import cKDTree

    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(cells, dtype=np.float64).reshape(-1, 3)
    if rows.shape[0] != cells.shape[0]:
        raise InitialStateError("E0510", f"the initial state has {rows.shape[0]} rows for {cells.shape[0]} fluid cells")
    tol = TOLERANCE * box_extent
    reach = tol * (1.0 + 1e-9)  # candidates by the tree, decided by _near
    close = cKDTree(cells).query_pairs(reach, output_type="ndarray")
    if close.size:
        close = close[_near(cells[close[:, 0]], cells[close[:, 1]], tol)]
        if close.size:
            a, b = sorted(close[np.lexsort((close[:, 1], close[:, 0]))][0])
            raise InitialStateError("E0511", f"cells {a} and {b} lie within the pairing tolerance of each other")
    dist, idx = cKDTree(rows).query(cells, k=2, distance_upper_bound=reach)
    hit = np.isfinite(dist)
    for j in range(2):
        ok = hit[:, j]
        hit[ok, j] = _near(cells[ok], rows[idx[ok, j]], tol)
    count = hit.sum(axis=1)
    bad = np.flatnonzero(count != 1)
    if bad.size:
        c = int(bad[0])
        raise InitialStateError("E0511", f"cell {c} matches {'no' if count[c] == 0 else 'several'} rows of the initial state")
    match = np.where(hit[:, 0], idx[:, 0], idx[:, 1]).astype(np.int64)
    order = np.argsort(match, kind="stable")
    twice = np.flatnonzero(match[order][1:] == match[order][:-1])
    if twice.size:
        a, b = int(order[twice[0]]), int(order[twice[0] + 1])
        raise InitialStateError("E0511", f"cells {a} and {b} match one row of the initial state")
    return match


def conserved(values: np.ndarray, gamma: float) -> np.ndarray:
    """5 x n conserved state of n x 5 rows (density, velocity, pressure), operation for operation
    as EulerPhysics::set_cell (physics/src/euler_physics.cpp:812-816)."""
    w = np.asarray(values, dtype=np.float64)
    rho, u, v, ww, p = (w[:, k] for k in range(5))
    return np.stack([rho, rho * u, rho * v, rho * ww, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + ww * ww)])


def start_state(state: InitialState, cell_centroids: np.ndarray, points: np.ndarray, gamma: float) -> np.ndarray:
    """The conserved start state (5 x n, FP64, flat) of the mesh's cells from a read file: rows
    paired with the cells by centroid, then set_cell's formulas."""
    rows = pair(state.centroids, cell_centroids, extent(points))
    return conserved(state.values[rows], gamma).reshape(-1)
