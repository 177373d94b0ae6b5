"""The oracle's Jacobian blocks (dual numbers through its own bodies: face_jacobian,
boundary_jacobian, dwdu) against what the mathematics fixes: central differences of the same
bodies' values, two steps to tell a smooth point from a switch of the flux (a column whose two
differences disagree straddles a kink and is not checked). This pins the oracle's Jacobian, against
which the device's is checked face by face."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from oracle import oracle as orc

GAMMA, RGAS = 1.4, 287.0
W_FREE = np.array([1.0, 0.4, 0.1, -0.05, 1.0 / GAMMA])
PREF = W_FREE[4]  # primitives carry the gauge pressure p - p_ref


def states(rng, count):
    """Subsonic face states with the gauge pressure."""
    rho = rng.uniform(0.6, 1.4, count)
    vel = rng.normal(0.0, 0.25, (count, 3)) + W_FREE[1:4]
    p = rng.uniform(0.6, 1.4, count) / GAMMA
    return np.column_stack([rho, vel, p - PREF])


def normals(rng, count):
    n = rng.normal(size=(count, 3))
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def central(fun, w, j, h):
    e = np.zeros(5)
    e[j] = h
    return (fun(w + e) - fun(w - e)) / (2.0 * h)


def check_columns(got, fun, w, tol=2e-6):
    scale = max(np.abs(got).max(), 1e-300)
    checked = 0
    for j in range(5):
        h = 1e-6 * max(1.0, abs(w[j]))
        a, b = central(fun, w, j, h), central(fun, w, j, 0.25 * h)
        if np.abs(a - b).max() > 1e-5 * scale:
            continue
        checked += 1
        assert np.abs(got[:, j] - b).max() <= tol * scale, (j, got[:, j], b)
    return checked


@pytest.mark.parametrize("flux", [0, 1, 2])
def test_face_jacobians_are_the_flux_derivatives(flux):
    rng = np.random.default_rng(10 + flux)
    wl, wr, nn = states(rng, 150), states(rng, 150), normals(rng, 150)
    checked = 0
    for a, b, n in zip(wl, wr, nn):
        jac = orc.face_jacobian(flux, a, b, n, GAMMA, 0.4, PREF)
        checked += check_columns(jac[0], lambda w: orc.flux(flux, w, b, n, GAMMA, 0.4, PREF), a)
        checked += check_columns(jac[1], lambda w: orc.flux(flux, a, w, n, GAMMA, 0.4, PREF), b)
    assert checked >= 0.97 * 150 * 10


@pytest.mark.parametrize("flux", [0, 1, 2])
def test_boundary_jacobians_go_through_the_boundary_state(flux):
    rng = np.random.default_rng(20 + flux)
    kinds = np.array([meshes.STATE, meshes.FARFIELD, meshes.INLET, meshes.OUTLET, meshes.SLIP, meshes.WALL])
    data = cases.boundary_data(kinds, W_FREE)
    checked = total = 0
    for wf, n in zip(states(rng, 40), normals(rng, 40)):
        for k, kind in enumerate(kinds):
            got = orc.boundary_jacobian(flux, wf, kind, data[:, k], n, GAMMA, RGAS, 0.4, PREF)
            if kind in (meshes.SLIP, meshes.WALL):
                want = np.zeros((5, 5))
                want[1:4, 4] = n
                np.testing.assert_array_equal(got, want)
                continue

            def fb(w, k=k, kind=kind):
                wb = orc.boundary_value(w, kind, data[:, k], n, GAMMA, RGAS, PREF)
                return orc.flux(flux, w, wb, n, GAMMA, 0.4, PREF)

            checked += check_columns(got, fb, wf)
            total += 5
    assert checked >= 0.97 * total


def test_dwdu_is_the_primitives_derivative():
    rng = np.random.default_rng(3)
    ur = np.array([1.0, 0.4, 0.1, -0.05, 1.0 / (GAMMA - 1.0) / GAMMA + 0.5 * 0.1725])
    for _ in range(50):
        dq = rng.normal(0.0, 0.05, 5)
        check_columns(orc.dwdu(dq, ur, GAMMA), lambda q: orc.to_primitive(q, ur, GAMMA), dq, tol=1e-7)


def test_fp32_blocks_are_the_fp64_blocks_to_fp32_precision():
    rng = np.random.default_rng(4)
    for a, b, n in zip(states(rng, 30), states(rng, 30), normals(rng, 30)):
        for flux in (0, 1, 2):
            j64 = orc.face_jacobian(flux, a, b, n, GAMMA, 0.4, PREF, policy=0)
            j32 = orc.face_jacobian(flux, a, b, n, GAMMA, 0.4, PREF, policy=1)
            assert np.abs(j64 - j32).max() <= 1e-4 * np.abs(j64).max()
