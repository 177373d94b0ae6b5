"""Meshes and their host set-up (include/mxmesh.h) as numpy arrays.

The set-up mirrors what the reference computes on the host before its device stepper runs:
canonical face order and RCM renumbering (core/src/renumber.cpp:109-148), FP64 fan-triangle
geometry (core/src/geometry.cpp:34-198), pack_geometry (codegen/cpp/geometry.cpp:166-307), the face2
stencil (codegen/cpp/geometry.cpp:432-617) and Euler::compute_gradient_weights
(physics/src/euler.cpp:340-416).
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from autan_amd import _lib

# BoundaryKind codes (physics/include/autan/physics/euler.hpp:30-37).
STATE, FARFIELD, SLIP, INLET, OUTLET, WALL = range(6)

_ARRAYS = (
    "owner neighbour cell_face_offsets cell_faces face_zone shift face_node_offsets face_nodes x y z "
    "face_normal face_area face_area_vector r_left r_right face_moment cell_volume cell_centroid face_centroid "
    "stencil_offsets stencil stencil_weights"
).split()
_DTYPES = {0: np.int32, 1: np.int64, 2: np.float64}


@dataclass
class Mesh:
    """A prepared mesh: topology, packed FP64 geometry and the face2 stencil with its weights.

    Arrays are structure of arrays, component c of entity i at c * n + i, as the reference's
    datasets (physics/include/autan/physics/euler.hpp:96-98).
    """

    n_cells: int
    n_faces: int
    n_interior: int
    all_planar: bool
    arrays: dict[str, np.ndarray] = field(repr=False)
    kept_boundaries: int = 0

    @property
    def n_boundary(self) -> int:
        return self.n_faces - self.n_interior

    @property
    def n_stencil(self) -> int:
        return int(self.arrays["stencil"].size)

    def __getattr__(self, name: str) -> np.ndarray:
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def centroids(self) -> np.ndarray:
        """Cell centroids, shape (n_cells, 3)."""
        return self.arrays["cell_centroid"].reshape(3, self.n_cells).T.copy()

    def face_centroids(self) -> np.ndarray:
        return self.arrays["face_centroid"].reshape(3, self.n_faces).T.copy()


class _Handle:
    def __init__(self, ptr: ctypes.c_void_p) -> None:
        self.ptr = ptr

    def __del__(self) -> None:
        if getattr(self, "ptr", None):
            _lib.mesh().mxmesh_destroy(self.ptr)
            self.ptr = None


def _call(fn, *args) -> None:
    err = ctypes.create_string_buffer(1024)
    if fn(*args, err, ctypes.c_size_t(len(err))) != 0:
        raise RuntimeError(err.value.decode(errors="replace"))


def _build(create, *args, order: str | None = "rcm", kinds: np.ndarray | None = None) -> Mesh:
    """order: "rcm" (the reference's renumbering), "morton" (a Z-curve of the cell centres) or
    None (the generator's order)."""
    lib = _lib.mesh()
    ptr = ctypes.c_void_p()
    _call(create, *args, ctypes.byref(ptr))
    h = _Handle(ptr)
    if order == "rcm":
        _call(lib.mxmesh_renumber_rcm, h.ptr)
    elif order == "morton":
        _call(lib.mxmesh_renumber_morton, h.ptr)
    elif order is not None:
        raise ValueError(f"unknown cell order {order!r}")
    _call(lib.mxmesh_prepare, h.ptr)
    sizes = (ctypes.c_int64 * 8)()
    lib.mxmesh_sizes(h.ptr, sizes)
    kept = 0
    if kinds is not None:
        kept = _set_kinds(h, kinds, int(sizes[1] - sizes[2]))
    mesh = Mesh(
        n_cells=int(sizes[0]),
        n_faces=int(sizes[1]),
        n_interior=int(sizes[2]),
        all_planar=bool(sizes[6]),
        arrays=_fetch(h),
        kept_boundaries=kept,
    )
    # The native handle stays with the mesh so that its boundary kinds can change later.
    mesh.__dict__["_handle"] = h
    return mesh


def _set_kinds(h: _Handle, kinds: np.ndarray, nb: int) -> int:
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    if k.size != nb:
        raise ValueError(f"{k.size} boundary kinds for {nb} boundary faces")
    kept = ctypes.c_int64()
    _call(_lib.mesh().mxmesh_set_boundary_kinds, h.ptr, k.ctypes.data_as(ctypes.c_void_p), ctypes.byref(kept))
    return int(kept.value)


def _fetch(h: _Handle) -> dict[str, np.ndarray]:
    lib = _lib.mesh()
    out = {}
    for name in _ARRAYS:
        data = ctypes.c_void_p()
        count = ctypes.c_int64()
        dtype = ctypes.c_int()
        _call(lib.mxmesh_array, h.ptr, name.encode(), ctypes.byref(data), ctypes.byref(count), ctypes.byref(dtype))
        dt = np.dtype(_DTYPES[dtype.value])
        n = int(count.value)
        if n == 0:
            out[name] = np.zeros(0, dtype=dt)
            continue
        # A view of the native array: valid while the handle (kept with the Mesh) lives.
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(data.value)
        out[name] = np.frombuffer(buf, dtype=dt)
    return out


def set_boundary_kinds(mesh: Mesh, kinds: np.ndarray) -> None:
    """Euler::set_boundary's weights for these condition codes (one per boundary face)."""
    h = mesh.__dict__["_handle"]
    mesh.kept_boundaries = _set_kinds(h, kinds, mesh.n_boundary)
    mesh.arrays["stencil_weights"] = _fetch(h)["stencil_weights"]


def periodic_hex(n: int, lo: float = 0.0, hi: float = 1.0, order: str | None = "rcm") -> Mesh:
    lib = _lib.mesh()
    return _build(lib.mxmesh_periodic_hex, ctypes.c_int(n), ctypes.c_double(lo), ctypes.c_double(hi), order=order)


def kuhn_tets(n: int, jitter: float = 0.1, seed: int = 0, order: str | None = "rcm", kinds=None) -> Mesh:
    lib = _lib.mesh()
    return _build(
        lib.mxmesh_kuhn_tets, ctypes.c_int(n), ctypes.c_double(jitter), ctypes.c_uint64(seed), order=order, kinds=kinds
    )


def bump_hex(nx, ny, nz, lx=3.0, ly=1.0, lz=1.0, h=0.1, w=0.2, x0=1.5, order="rcm", kinds=None) -> Mesh:
    lib = _lib.mesh()
    args = [ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz)]
    args += [ctypes.c_double(v) for v in (lx, ly, lz, h, w, x0)]
    return _build(lib.mxmesh_bump_hex, *args, order=order, kinds=kinds)


def bump_hybrid(nx, ny, nz, lx=3.0, ly=1.0, lz=1.0, h=0.1, w=0.2, x0=1.5, prism_layers=2, band=None,
                order="morton", kinds=None) -> Mesh:
    """The bump channel with a prism layer on the lower wall and a band of merged polyhedra."""
    lib = _lib.mesh()
    if band is None:
        lo = max(prism_layers, ny // 4)
        band = (lo, lo + 2 * ((ny - lo) // 4))
    lo, hi = band
    args = [ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz)]
    args += [ctypes.c_double(v) for v in (lx, ly, lz, h, w, x0)]
    args += [ctypes.c_int(prism_layers), ctypes.c_int(lo), ctypes.c_int(hi)]
    return _build(lib.mxmesh_bump_hybrid, *args, order=order, kinds=kinds)


def read_gmsh(path: str | os.PathLike, order: str | None = None, kinds=None) -> Mesh:
    """A Gmsh MSH 4.1 file (ASCII or binary); boundary zone k is the k-th 2-D physical group by tag."""
    lib = _lib.mesh()
    return _build(lib.mxmesh_read_gmsh, os.fsencode(str(path)), order=order, kinds=kinds)


def read_polymesh(path: str | os.PathLike, order: str | None = None, kinds=None) -> Mesh:
    lib = _lib.mesh()
    return _build(lib.mxmesh_read_polymesh, os.fsencode(str(path)), order=order, kinds=kinds)
