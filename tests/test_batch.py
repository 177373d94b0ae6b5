"""Batches (autan_amd.batch): the union of several meshes keeps the mesh contract, and a residual
on the union is each case's residual on its own (checked with the oracle, one scheme for all)."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import batch
from autan_amd.mesh import Mesh
from oracle.oracle import Oracle


def three_cases():
    return [cases.bump_case(12, 4, 2), cases.bump_case(16, 12, 8, hybrid=True), cases.bump_case(24, 8, 4)]


def as_mesh(u: batch.Union) -> Mesh:
    return Mesh(u.n_cells, u.n_faces, u.n_interior, u.all_planar, u.arrays)


def test_union_keeps_the_mesh_contract():
    cs = three_cases()
    u = batch.union([c.mesh for c in cs])
    assert np.all(u.case_first % batch.BLOCK == 0) and u.case_first[-1] == u.n_cells
    m = as_mesh(u)
    ni = m.n_interior
    own, nei = m.owner, m.neighbour
    assert np.all(own[:ni] < nei)
    assert np.all(np.diff(own[:ni].astype(np.int64) * m.n_cells + nei) >= 0)
    assert np.all(np.diff(m.face_zone.astype(np.int64) * m.n_cells + own[ni:]) >= 0)
    off, cf = m.cell_face_offsets, m.cell_faces
    for c in range(m.n_cells):
        faces = cf[off[c] : off[c + 1]]
        assert np.all(np.diff(faces) > 0)
        assert np.all((own[faces] == c) | ((faces < ni) & (nei[np.minimum(faces, ni - 1)] == c)))
    # filler cells: no faces, no stencil
    for j, c in enumerate(cs):
        c0 = int(u.case_first[j]) + c.mesh.n_cells
        c1 = int(u.case_first[j + 1])
        assert np.all(np.diff(off[c0 : c1 + 1]) == 0)
        assert np.all(np.diff(m.stencil_offsets[c0 : c1 + 1]) == 0)
        np.testing.assert_array_equal(m.cell_volume[c0:c1], 1.0)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_union_residual_is_each_cases_residual(policy):
    cs = three_cases()
    u = batch.union([c.mesh for c in cs])
    m = as_mesh(u)
    sc = cs[0].scheme
    sc.planar = False
    kinds = np.concatenate([c.bc_kind for c in cs])
    bc = np.concatenate([np.asarray(c.bc).reshape(5, -1) for c in cs], axis=1)
    ref = np.array(sc.reference)
    state = batch.union_state(u, [c.u for c in cs], [ref] * len(cs))
    whole = Oracle(m, sc, policy, kinds, bc.reshape(-1), None).residual(state).reshape(5, -1)
    for j, c in enumerate(cs):
        c.scheme.planar = False
        own = Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc, None).residual(c.u).reshape(5, -1)
        c0 = int(u.case_first[j])
        np.testing.assert_array_equal(whole[:, c0 : c0 + c.mesh.n_cells], own)
        np.testing.assert_array_equal(whole[:, c0 + c.mesh.n_cells : int(u.case_first[j + 1])], 0.0)


def test_split_and_union_state_round_trip():
    cs = three_cases()
    u = batch.union([c.mesh for c in cs])
    s = batch.union_state(u, [c.u for c in cs], [np.arange(5.0)] * 3)
    back = batch.split_state(u, s)
    for a, c in zip(back, cs):
        np.testing.assert_array_equal(a, c.u)
