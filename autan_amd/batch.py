"""Batches: several meshes (cases) stepped together by one device stepper.

The cases are joined into one index space (`union`): case k's cells start on a 256-cell block,
followed by filler cells (no faces, no stencil) up to the next block; its interior faces follow the
previous cases' interior faces and its boundary faces the previous cases' boundary faces, so the
union keeps the mesh contract (interior faces first by (owner, neighbour), boundary faces by
(zone, owner), each cell's faces ascending). Every kernel launch then covers every case; each
block belongs to one case and reads that case's parameters. Results per case are those of the
case stepped alone.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from autan_amd import _lib
from autan_amd import steady as _steady
from autan_amd.mesh import Mesh
from autan_amd.stepper import POLICIES, Scheme, _check, _d, _err, _Geometry, _i32, _p, _Stencil, _Topology

BLOCK = 256


@dataclass
class Case:
    mesh: Mesh
    scheme: Scheme
    u: np.ndarray  # 5 x n conserved
    bc_kind: np.ndarray
    bc: np.ndarray  # 5 x nb
    source: np.ndarray | None = None


@dataclass
class Union:
    arrays: dict
    n_cells: int
    n_faces: int
    n_interior: int
    all_planar: bool
    case_first: np.ndarray  # n_cases + 1
    case_cells: np.ndarray  # n_cases
    face_first: list  # per case: (first interior face, first boundary face) in the union


def union(meshes: list[Mesh]) -> Union:
    """The cases' meshes in one index space (see the module's docstring)."""
    k = len(meshes)
    ranges = [((m.n_cells + BLOCK - 1) // BLOCK) * BLOCK for m in meshes]
    first = np.concatenate([[0], np.cumsum(ranges)]).astype(np.int64)
    nc = int(first[-1])
    ni = [m.n_interior for m in meshes]
    nb = [m.n_boundary for m in meshes]
    NI = int(sum(ni))
    fi = np.concatenate([[0], np.cumsum(ni)]).astype(np.int64)
    fb = NI + np.concatenate([[0], np.cumsum(nb)]).astype(np.int64)
    nf = NI + int(sum(nb))

    def fmap(j, f):
        f = np.asarray(f, dtype=np.int64)
        return np.where(f < ni[j], fi[j] + f, fb[j] + (f - ni[j]))

    order = []  # union face -> (case, local face)
    for j, m in enumerate(meshes):
        order.append((j, np.arange(ni[j])))
    for j, m in enumerate(meshes):
        order.append((j, np.arange(ni[j], m.n_faces)))

    def face_soa(name, comps):
        out = np.zeros((comps, nf))
        at = 0
        for j, loc in order:
            a = meshes[j].arrays[name].reshape(comps, meshes[j].n_faces)
            out[:, at : at + loc.size] = a[:, loc]
            at += loc.size
        return out.reshape(-1)

    owner = np.zeros(nf, dtype=np.int32)
    at = 0
    for j, loc in order:
        owner[at : at + loc.size] = meshes[j].owner[loc] + first[j]
        at += loc.size
    neighbour = np.concatenate([m.neighbour.astype(np.int64) + first[j] for j, m in enumerate(meshes)]).astype(np.int32)
    max_zone = max(int(m.face_zone.max()) + 1 if m.n_boundary else 1 for m in meshes)
    face_zone = np.concatenate([m.face_zone.astype(np.int64) + j * max_zone for j, m in enumerate(meshes)]).astype(np.int32)

    cfo = np.zeros(nc + 1, dtype=np.int64)
    cf = []
    so = np.zeros(nc + 1, dtype=np.int64)
    st = []
    sw = [[], [], []]
    vol = np.ones(nc)
    cc = np.zeros((3, nc))
    for j, m in enumerate(meshes):
        c0, n = int(first[j]), m.n_cells
        counts = np.diff(m.cell_face_offsets)
        cfo[c0 + 1 : c0 + n + 1] = counts
        cf.append(fmap(j, m.cell_faces))
        scounts = np.diff(m.stencil_offsets)
        so[c0 + 1 : c0 + n + 1] = scounts
        e = m.stencil.astype(np.int64)
        st.append(np.where(e >= 0, e + c0, -(fmap(j, np.maximum(-e - 1, 0)) + 1)))
        w = m.stencil_weights.reshape(3, -1)
        for q in range(3):
            sw[q].append(w[q])
        vol[c0 : c0 + n] = m.cell_volume
        cc[:, c0 : c0 + n] = m.cell_centroid.reshape(3, n)
    cfo = np.cumsum(cfo)
    so = np.cumsum(so)
    arrays = {
        "owner": owner,
        "neighbour": neighbour,
        "cell_face_offsets": cfo,
        "cell_faces": np.concatenate(cf).astype(np.int32),
        "face_zone": face_zone,
        "face_normal": face_soa("face_normal", 3),
        "face_area": face_soa("face_area", 1),
        "face_area_vector": face_soa("face_area_vector", 3),
        "r_left": face_soa("r_left", 3),
        "r_right": face_soa("r_right", 3),
        "face_moment": face_soa("face_moment", 9),
        "face_centroid": face_soa("face_centroid", 3),
        "cell_volume": vol,
        "cell_centroid": cc.reshape(-1),
        "stencil_offsets": so,
        "stencil": np.concatenate(st).astype(np.int32),
        "stencil_weights": np.concatenate([np.concatenate(sw[q]) for q in range(3)]),
    }
    return Union(
        arrays, nc, nf, NI, all(m.all_planar for m in meshes), first,
        np.array([m.n_cells for m in meshes], dtype=np.int64), [(int(fi[j]), int(fb[j])) for j in range(k)],
    )


def union_state(u: Union, states: list[np.ndarray], fill: list[np.ndarray]) -> np.ndarray:
    """The cases' 5 x n states in the union, filler cells holding `fill[k]` (5 values)."""
    out = np.zeros((5, u.n_cells))
    for j, s in enumerate(states):
        c0, c1, n = int(u.case_first[j]), int(u.case_first[j + 1]), int(u.case_cells[j])
        out[:, c0 : c0 + n] = np.asarray(s, dtype=np.float64).reshape(5, n)
        out[:, c0 + n : c1] = np.asarray(fill[j], dtype=np.float64)[:, None]
    return out.reshape(-1)


def split_state(u: Union, a: np.ndarray) -> list[np.ndarray]:
    a = np.asarray(a).reshape(5, u.n_cells)
    return [a[:, int(u.case_first[j]) : int(u.case_first[j]) + int(u.case_cells[j])].reshape(-1).copy()
            for j in range(len(u.case_cells))]


class BatchStepper:
    """mxeuler_create_batch over the cases (one policy, one flux and order for all)."""

    def __init__(self, cases: list[Case], policy: str = "fp32+acc64", stream: int | None = None) -> None:
        self.lib = _lib.euler()
        self.cases = cases
        self.policy = POLICIES[policy] if isinstance(policy, str) else int(policy)
        meshes = [c.mesh for c in cases]
        self.u = union(meshes)
        u = self.u
        self.n = len(cases)
        kinds = np.concatenate([np.asarray(c.bc_kind, dtype=np.float64).reshape(-1) for c in cases])
        nb = [m.n_boundary for m in meshes]
        bc = np.concatenate([np.asarray(c.bc, dtype=np.float64).reshape(5, n) for c, n in zip(cases, nb)], axis=1)
        src = None
        if any(c.source is not None for c in cases):
            src = union_state(u, [np.zeros(5 * c.mesh.n_cells) if c.source is None else c.source for c in cases],
                              [np.zeros(5)] * self.n)
        planar = u.all_planar
        schemes = (type(cases[0].scheme.native(planar)) * self.n)(*[c.scheme.native(planar) for c in cases])
        keep = []

        def a(x):
            x = np.ascontiguousarray(x)
            keep.append(x)
            return x.ctypes.data

        A = u.arrays
        topo = _Topology(u.n_cells, u.n_faces, u.n_interior, a(A["owner"]), a(A["neighbour"]),
                         a(A["cell_face_offsets"]), a(A["cell_faces"]))
        geo = _Geometry(a(A["face_normal"]), a(A["face_area"]), a(A["r_left"]), a(A["r_right"]), a(A["face_moment"]),
                        a(A["cell_volume"]), int(planar))
        st = _Stencil(a(A["stencil_offsets"]), a(A["stencil"]), a(A["stencil_weights"]))
        first = np.ascontiguousarray(u.case_first, dtype=np.int64)
        cells = np.ascontiguousarray(u.case_cells, dtype=np.int64)
        kinds = np.ascontiguousarray(kinds)
        bc = np.ascontiguousarray(bc.reshape(-1))
        handle = ctypes.c_void_p()
        err = _err()
        rc = self.lib.mxeuler_create_batch(
            _i32(self.policy), _i32(self.n), schemes, _p(first.ctypes.data), _p(cells.ctypes.data), ctypes.byref(topo),
            ctypes.byref(geo), ctypes.byref(st), _p(kinds.ctypes.data), _p(bc.ctypes.data),
            _p(None if src is None else src.ctypes.data), _p(stream), ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        self.handle = handle
        self.fill = [np.array(c.scheme.reference, dtype=np.float64) for c in cases]

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.mxeuler_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        self.close()

    def _call(self, fn, *args) -> None:
        err = _err()
        _check(fn(self.handle, *args, err, ctypes.c_size_t(len(err))), err)

    def upload(self, states: list[np.ndarray] | None = None) -> None:
        states = states if states is not None else [c.u for c in self.cases]
        x = np.ascontiguousarray(union_state(self.u, states, self.fill))
        self._call(self.lib.mxeuler_upload, _p(x.ctypes.data))

    def download(self) -> list[np.ndarray]:
        x = np.empty(5 * self.u.n_cells)
        self._call(self.lib.mxeuler_download, _p(x.ctypes.data))
        return split_state(self.u, x)

    def step(self, cfl: float, local: bool = False):
        r = np.empty(5 * self.n)
        h = np.empty(self.n)
        inv = np.empty(self.n, dtype=np.int32)
        self._call(self.lib.mxeuler_batch_step, _d(cfl), _i32(local), _p(r.ctypes.data), _p(h.ctypes.data),
                   _p(inv.ctypes.data))
        return r.reshape(self.n, 5), h, [None if v < 0 else int(v) for v in inv]

    def loads(self, k: int, faces: np.ndarray, center=(0.0, 0.0, 0.0), p_inf: float = 0.0):
        """EulerPhysics::loads over case k's boundary faces `faces` (the case's own face indices)."""
        m = self.cases[k].mesh
        fi, fb = self.u.face_first[k]
        f = np.asarray(faces, dtype=np.int64)
        g = np.ascontiguousarray(np.where(f < m.n_interior, fi + f, fb + (f - m.n_interior)), dtype=np.int32)
        av = np.ascontiguousarray(m.face_area_vector.reshape(3, -1)[:, f])
        fc = np.ascontiguousarray(m.face_centroid.reshape(3, -1)[:, f])
        c = np.ascontiguousarray(center, dtype=np.float64)
        force = (ctypes.c_double * 3)()
        moment = (ctypes.c_double * 3)()
        self._call(self.lib.mxeuler_loads, _p(g.ctypes.data), ctypes.c_int64(g.size), _p(av.ctypes.data),
                   _p(fc.ctypes.data), _p(c.ctypes.data), _d(p_inf), force, moment)
        return np.array(force[:]), np.array(moment[:])

    def tune(self, key: str, value: int) -> None:
        self._call(self.lib.mxeuler_tune, ctypes.c_char_p(key.encode()), _i32(value))

    def set_flags(self, flags) -> None:
        f = np.ascontiguousarray(flags, dtype=np.int32)
        self._call(self.lib.mxeuler_batch_flags, _p(f.ctypes.data))

    def autotune(self, steps: int = 5, cfl: float = _steady.STEADY_CFL, local: bool = True) -> str:
        """Keeps the fastest kernel variant (same bits) for these cases; the state is restored."""
        report = ctypes.create_string_buffer(2048)
        self._call(self.lib.mxeuler_autotune, _i32(steps), _d(cfl), _i32(local), report, ctypes.c_size_t(len(report)))
        return report.value.decode()

    def steady(self, max_iterations: int = 10000, residual_orders: float = 8.0, cfl: float = _steady.STEADY_CFL,
               level_rule: bool | None = None, start_order: str = "first", scales=None, references=None,
               history: bool = True):
        """Every case's steady run (autan_amd.steady.solve's rules per case); scales and references:
        5 per case (references: the residuals to drop from, else each case's iteration 1)."""
        if level_rule is None:
            level_rule = self.policy == 1
        keep: list = []
        opts = _steady._options(max_iterations, residual_orders, cfl, level_rule, start_order, None, None, keep)
        cs = None if scales is None else np.ascontiguousarray(scales, dtype=np.float64).reshape(self.n, 5)
        cr = None if references is None else np.ascontiguousarray(references, dtype=np.float64).reshape(self.n, 5)
        res = (_steady._Result * self.n)()
        hist = np.zeros(5 * max_iterations * self.n) if history else None
        self._call(self.lib.mxeuler_batch_steady, ctypes.byref(opts), _p(None if cs is None else cs.ctypes.data),
                   _p(None if cr is None else cr.ctypes.data), res, _p(None if hist is None else hist.ctypes.data))
        out = []
        for j in range(self.n):
            h = hist[5 * max_iterations * j : 5 * max_iterations * (j + 1)] if hist is not None else np.zeros(0)
            out.append(_steady._finish(res[j], h, np.zeros(max_iterations, dtype=np.int32)))
        return out
