"""The reference's Ringleb fixture (tests/data/mesh/ringleb/ringleb-n004.{msh,states.json}, copied
under tests/golden/meshes/ringleb): the exact Ringleb-flow states (gamma 1.4, stagnation density
and speed of sound 1) at the centroids of the boundary faces of the reference's own Gmsh mesh,
subsonic and supersonic (Mach 0.51 to 2.03).

- The Gmsh reader and the fan-triangle geometry put every boundary face where the fixture has it
  (zones by physical group, centroids to round-off).
- The oracle's physical flux carries the total enthalpy: F_E / F_rho = H = c0^2 / (gamma - 1) = 2.5
  for every exact state (homenthalpic flow), with and without a gauge pressure.
- The characteristic boundary states reproduce an exact state given as both the interior and the
  outside state: farfield with the mesh's normal (subsonic normal flow), along the flow
  (supersonic outflow where the state is supersonic) and against it (supersonic inflow); outlet at
  the state's own pressure.
- Every flux of the exact state against its own boundary state is the physical flux."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from autan_amd import mesh as meshes
from oracle import oracle as orc

HERE = Path(__file__).resolve().parent / "golden" / "meshes" / "ringleb"
GAMMA = 1.4
RGAS = 1.0 / GAMMA
ZONES = ["k-min", "k-max", "inflow", "outflow"]


@pytest.fixture(scope="module")
def ringleb():
    m = meshes.read_gmsh(HERE / "ringleb-n004.msh", order=None)
    doc = json.loads((HERE / "ringleb-n004.states.json").read_text())
    assert doc["schema"] == "autan.face-states/1" and doc["oracle"] == "ringleb"
    fc = m.face_centroids()[m.n_interior :]
    normals = np.asarray(m.face_normal).reshape(3, -1)[:, m.n_interior :].T
    faces = []
    for name in ZONES:
        z = doc["zones"][name]
        c = np.asarray(z["centroids"])
        d = np.linalg.norm(fc[:, None, :] - c[None, :, :], axis=2)
        at = np.argmin(d, axis=0)
        faces.append((name, at, c, np.asarray(z["values"]), d[at, np.arange(c.shape[0])]))
    return m, faces, normals


def test_boundary_faces_sit_where_the_fixture_has_them(ringleb):
    m, faces, _ = ringleb
    for zone, (name, at, c, values, dist) in enumerate(faces):
        assert len(set(at.tolist())) == len(at), name
        assert np.all(m.face_zone[at] == zone), name
        assert dist.max() <= 1e-14, name


@pytest.mark.parametrize("pref", [0.0, 0.37])
def test_the_physical_flux_carries_the_total_enthalpy(ringleb, pref):
    _, faces, normals = ringleb
    for name, at, _, values, _ in faces:
        for w, n in zip(values, normals[at]):
            wg = w.copy()
            wg[4] -= pref  # the gauge pressure the bodies take
            f = orc.physical_flux(wg, n, GAMMA, pref)
            speed = np.linalg.norm(w[1:4])
            mass = w[0] * (w[1:4] @ n)
            assert abs(f[0] - mass) <= 4e-16 * w[0] * speed
            if abs(mass) > 1e-3 * w[0] * speed:
                assert f[4] / f[0] == pytest.approx(1.0 / (GAMMA - 1.0), rel=1e-13), name
            # momentum: rho v (v . n) + (p - p_ref) n, the gauge pressure
            want = w[0] * w[1:4] * (w[1:4] @ n) + wg[4] * n
            assert np.abs(f[1:4] - want).max() <= 1e-15 * (w[0] * speed * speed + w[4]), name


def directions(w, n):
    v = w[1:4]
    s = v / np.linalg.norm(v)
    return {"mesh normal": n, "along the flow": s, "against the flow": -s}


def test_characteristic_states_reproduce_an_exact_state(ringleb):
    _, faces, normals = ringleb
    branches = set()
    for name, at, _, values, _ in faces:
        for w, n in zip(values, normals[at]):
            c = np.sqrt(GAMMA * w[4] / w[0])
            for label, d in directions(w, n).items():
                vn = w[1:4] @ d
                branches.add("out" if vn >= c else ("in" if vn <= -c else "subsonic"))
                got = orc.boundary_value(w, meshes.FARFIELD, w, d, GAMMA, RGAS, 0.0)
                np.testing.assert_allclose(got, w, rtol=2e-14, atol=1e-15, err_msg=f"{name} {label}")
            data = np.array([0.0, 0.0, 0.0, 0.0, w[4]])
            got = orc.boundary_value(w, meshes.OUTLET, data, n, GAMMA, RGAS, 0.0)
            np.testing.assert_allclose(got, w, rtol=1e-15, atol=1e-16, err_msg=f"{name} outlet")
    # The fixture reaches all three of the farfield's branches.
    assert branches == {"out", "in", "subsonic"}


@pytest.mark.parametrize("flux", [0, 1, 2])
def test_each_flux_against_the_reproduced_state_is_the_physical_flux(ringleb, flux):
    _, faces, normals = ringleb
    for name, at, _, values, _ in faces:
        for w, n in zip(values, normals[at]):
            wb = orc.boundary_value(w, meshes.FARFIELD, w, n, GAMMA, RGAS, 0.0)
            got = orc.flux(flux, w, wb, n, GAMMA, 0.5, 0.0)
            want = orc.physical_flux(w, n, GAMMA, 0.0)
            scale = np.abs(want).max() + w[4]
            assert np.abs(got - want).max() <= 1e-13 * scale, (name, flux)
