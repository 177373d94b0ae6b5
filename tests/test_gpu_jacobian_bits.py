"""The device's Jacobian blocks (mxeuler_residual_and_jacobian) against the oracle's (the same flux,
reconstruction and boundary bodies on the oracle's own dual numbers, pinned by central differences
in test_oracle_jacobian.py), face by face and in the policy's compute type: bit for bit, except at
farfield and inlet faces, whose boundary states call exp / log (the device's and glibc's differ in
the last bits), which are checked within 64 units of the compute type."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd.stepper import FLUXES, DeviceStepper
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def p_ref(sc) -> float:
    r = sc.reference
    return (sc.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0])


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name,flux,second", [("tets", "hllc", True), ("hex", "roe", True), ("hex", "ausm+up", False)])
def test_device_blocks_are_the_oracles(name, flux, second, policy):
    c = cases.tets_case(4, flux=flux, second_order=second) if name == "tets" else \
        cases.fixture_case("hex", flux=flux, second_order=second)
    m, sc = c.mesh, c.scheme
    t = np.float64 if policy == "fp64" else np.float32
    pol = 0 if policy == "fp64" else 1
    print(f"jacobian bits {name} {flux} second={second} {policy}", flush=True)
    dev = DeviceStepper(m, sc, policy, c.bc_kind, c.bc, c.source)
    parts = dev.residual_and_jacobian(c.u)
    arr = dev.arrays()
    dev.close()
    nc, ni, nb = m.n_cells, m.n_interior, m.n_boundary
    prim, grad = arr["prim"].reshape(5, nc).astype(t), arr["grad"].reshape(15, nc).astype(t)
    phi = arr["phi"].reshape(5, nc).astype(t)
    nrm = m.face_normal.reshape(3, -1).astype(t)
    rl, rr = m.r_left.reshape(3, -1).astype(t), m.r_right.reshape(3, -1).astype(t)
    area = m.face_area.astype(t)
    g, pref, minf, rgas = t(sc.gamma), t(p_ref(sc)), t(sc.mach_ref), t(sc.gas_constant)
    bc = np.asarray(c.bc, dtype=np.float64).reshape(5, nb).astype(t)
    kinds = np.asarray(c.bc_kind, dtype=np.int64)
    kind = FLUXES[flux]

    def side(cell, r):
        w = prim[:, cell]
        if not second:
            return w
        gg, lim = grad[:, cell], phi[:, cell]
        d = np.array([gg[3 * k] * r[0] + gg[3 * k + 1] * r[1] + gg[3 * k + 2] * r[2] for k in range(5)], dtype=t)
        f = (w + lim * d).astype(t)
        return f if (f[0] > t(0) and f[4] + pref > t(0)) else w

    def widen(x):
        return np.asarray(x, dtype=np.float64)

    for f in range(ni):
        wl, wr = side(m.owner[f], rl[:, f]), side(m.neighbour[f], rr[:, f])
        want = orc.face_jacobian(kind, widen(wl), widen(wr), widen(nrm[:, f]), float(g), float(minf), float(pref),
                                 policy=pol).astype(t)
        assert np.array_equal(parts["face_jac"][:, :, :, f], widen(area[f] * want)), f
    eps = np.finfo(t).eps
    for b in range(nb):
        f = ni + b
        wf = side(m.owner[f], rl[:, f])
        want = orc.boundary_jacobian(kind, widen(wf), kinds[b], widen(bc[:, b]), widen(nrm[:, f]), float(g), float(rgas),
                                     float(minf), float(pref), policy=pol).astype(t)
        got = parts["bnd_jac"][:, :, b]
        if kinds[b] in (meshes.FARFIELD, meshes.INLET):
            assert np.abs(got - widen(area[f] * want)).max() <= 64 * eps * np.abs(got).max(), b
        else:
            assert np.array_equal(got, widen(area[f] * want)), b
