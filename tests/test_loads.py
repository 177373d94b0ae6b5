"""Wall loads (Euler::boundary_pressure, EulerPhysics::loads): the oracle's face force against
what the mathematics fixes (physics/src/euler.cpp:1187-1223; physics/src/euler_physics.cpp:527-562,
930-948)."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd.stepper import freestream_scheme
from oracle import oracle as orc
from oracle.oracle import Oracle

GAMMA = 1.4


def linear_pressure_case(policy_second=True):
    """The unit-cube hex fixture at rest with p = p0 + a x + b y + c z and every boundary face a
    `state` face holding the exact value: the unlimited least-squares gradient is exact, so the
    face pressure is the field at the face centroid."""
    m = meshes.read_polymesh(cases.GOLDEN / "meshes" / "hex")
    meshes.set_boundary_kinds(m, np.zeros(m.n_boundary, dtype=np.int32))
    p0, a, b, c = 1.0 / GAMMA, 0.05, -0.03, 0.02

    def field(x):
        p = p0 + a * x[:, 0] + b * x[:, 1] + c * x[:, 2]
        return cases.conserved(np.ones(len(x)), 0 * p, 0 * p, 0 * p, p), p

    u, _ = field(m.centroids())
    xf = m.face_centroids()[m.n_interior :]
    _, pb = field(xf)
    bc = np.vstack([np.ones(len(xf)), 0 * pb, 0 * pb, 0 * pb, pb])
    sc = freestream_scheme(1.0, (0.0, 0.0, 0.0), p0, flux="roe", second_order=True, limiter=False)
    return m, sc, u.reshape(-1), bc, (p0, a, b, c)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_face_force_integrates_a_linear_pressure(policy):
    m, sc, u, bc, (p0, a, b, c) = linear_pressure_case()
    ora = Oracle(m, sc, policy, np.zeros(m.n_boundary), bc.reshape(-1), None)
    xf = m.face_centroids()
    faces = np.arange(m.n_interior, m.n_faces)
    force = ora.face_force(u, faces)
    tol = 1e-13 if policy == "fp64" else 3e-7
    # Each face: p(x_f) S_f, with S_f the outward area vector.
    s = m.face_normal.reshape(3, -1)[:, faces] * m.face_area[faces]
    p_f = p0 + a * xf[faces, 0] + b * xf[faces, 1] + c * xf[faces, 2]
    np.testing.assert_allclose(force, p_f * s, rtol=0, atol=tol)
    # The plane x = 0: -(p0 + b/2 + c/2) e_x over the unit square.
    on = np.abs(xf[faces, 0]) < 1e-12
    total = force[:, on].sum(axis=1)
    np.testing.assert_allclose(total, [-(p0 + b / 2 + c / 2), 0.0, 0.0], rtol=0, atol=10 * tol)
    # The closed surface: sum p S over the outward area vectors is grad p times the volume.
    np.testing.assert_allclose(force.sum(axis=1), [a, b, c], rtol=0, atol=30 * tol)


def test_face_force_falls_back_to_the_cell_pressure():
    """Where the reconstructed pressure is not positive the face takes the cell value (the
    reconstruct fallback): p = -0.1 + 1.2 x is positive at every cell centroid but negative on the
    plane x = 0."""
    m, sc, u, bc, _ = linear_pressure_case()
    x = m.centroids()
    xf = m.face_centroids()
    faces = np.arange(m.n_interior, m.n_faces)
    p_cell = -0.1 + 1.2 * x[:, 0]
    p_face = -0.1 + 1.2 * xf[faces, 0]
    assert p_cell.min() > 0 and p_face.min() < 0
    u = cases.conserved(np.ones(m.n_cells), 0 * p_cell, 0 * p_cell, 0 * p_cell, p_cell).reshape(-1)
    bc = np.vstack([np.ones(len(faces)), 0 * p_face, 0 * p_face, 0 * p_face, p_face])
    ora = Oracle(m, sc, "fp64", np.zeros(m.n_boundary), bc.reshape(-1), None)
    force = ora.face_force(u, faces)
    s = m.face_normal.reshape(3, -1)[:, faces] * m.face_area[faces]
    low = p_face < 0
    own = m.owner[faces]
    np.testing.assert_allclose(force[:, low], (p_cell[own] * s)[:, low], rtol=1e-13)
    np.testing.assert_allclose(force[:, ~low], (p_face * s)[:, ~low], rtol=1e-12, atol=1e-15)


def test_exact_loads_sum():
    """loads' sums are exact sums: order-independent to the bit."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000) * 10.0 ** rng.integers(-5, 5, 1000)
    assert orc.exact_sum(x) == orc.exact_sum(x[::-1]) == orc.exact_sum(rng.permutation(x))


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name", ["bump-hybrid", "tets", "poly"])
def test_conservation_holds_to_round_off(name, policy):
    """Sum_c R^F_c = sum_b F_b where every interior flux enters its two cells with opposite signs:
    E is the rounding of the per-cell FP64 sums, at most (faces per cell) u D, u = 2^-53; the sums
    of the residual are those of the oracle's own residual without source."""
    case = {
        "bump-hybrid": lambda: cases.bump_case(16, 12, 8, hybrid=True),
        "tets": lambda: cases.tets_case(5),
        "poly": lambda: cases.fixture_case("poly", "hllc"),
    }[name]()
    ora = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    r = ora.conservation(case.u)
    m = int(np.diff(case.mesh.cell_face_offsets).max())
    assert np.all(np.abs(r["E"]) <= m * 2.0**-53 * r["D"])
    assert np.all(r["D"] > 0)
    plain = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, None)
    res = plain.residual(case.u).reshape(5, -1)
    assert list(r["faces"]) == [orc.exact_sum(res[k]) for k in range(5)]
    flux = ora.face_fluxes().reshape(5, -1)[:, case.mesh.n_interior :]
    assert list(r["boundary"]) == [orc.exact_sum(flux[k]) for k in range(5)]
