"""The approximate Jacobian from the device (mxeuler_residual_and_jacobian) against what the
mathematics fixes: each face block is the derivative of that face's flux, which central differences
of the oracle's flux body give (the reconstruction's gradients and limiter held fixed, as the
reference's face_jac and bnd_jac define them, kernels/euler.py:854-860, 1051-1057); and at first
order with `state` boundaries (whose boundary values do not depend on the cells) the assembled
matrix is the residual's exact Jacobian, which central differences of the device residual give.
A boundary face's block is the derivative through boundary_value of the reconstructed state, as
the reference's boundary bodies compute it (kernels/euler.py:868-940); slip and wall faces take
the pressure flux alone there."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import jacobian
from autan_amd import mesh as meshes
from autan_amd.stepper import FLUXES, DeviceStepper
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def p_ref(sc) -> float:
    r = sc.reference
    return (sc.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0])


def reconstruct(w, g, lim, r, pref):
    """kernels/euler.py:619-627 on one face side: W + phi (grad W . r), the cell value where the
    density or pressure would not stay positive."""
    d = np.array([g[3 * k] * r[0] + g[3 * k + 1] * r[1] + g[3 * k + 2] * r[2] for k in range(5)])
    f = w + lim * d
    return f if (f[0] > 0.0 and f[4] + pref > 0.0) else w


def fd(fun, w, j, h):
    e = np.zeros(5)
    e[j] = h
    return (fun(w + e) - fun(w - e)) / (2.0 * h)


def check_block(got, fun, w, tol=2e-6):
    """got (5 x 5, dF/dW) against central differences of fun at w, with two steps: a column whose
    two differences disagree straddles a switch of the flux (a kink) and is not checked."""
    scale = max(np.abs(got).max(), 1e-300)
    checked = 0
    for j in range(5):
        h = 1e-6 * max(1.0, abs(w[j]))
        a, b = fd(fun, w, j, h), fd(fun, w, j, 0.25 * h)
        if np.abs(a - b).max() > 1e-5 * scale:
            continue
        checked += 1
        assert np.abs(got[:, j] - b).max() <= tol * scale, (j, got[:, j], b)
    return checked


CASES = [
    ("tets", "hllc", True), ("tets", "roe", True), ("tets", "ausm+up", True),
    ("hex", "hllc", False), ("hex", "roe", True), ("hex", "ausm+up", True),
]


def build(name, flux, second):
    if name == "tets":
        return cases.tets_case(4, flux=flux, second_order=second)
    return cases.fixture_case("hex", flux=flux, second_order=second)


@pytest.mark.parametrize("name,flux,second", CASES)
def test_face_blocks_are_the_flux_derivatives(name, flux, second):
    c = build(name, flux, second)
    m, sc = c.mesh, c.scheme
    print(f"jacobian faces {name} {flux} second={second}", flush=True)
    dev = DeviceStepper(m, sc, "fp64", c.bc_kind, c.bc, c.source)
    parts = dev.residual_and_jacobian(c.u)
    arr = dev.arrays()
    dev.close()
    pref = p_ref(sc)
    kind = FLUXES[flux]
    nrm = m.face_normal.reshape(3, -1)
    rl, rr = m.r_left.reshape(3, -1), m.r_right.reshape(3, -1)
    area = m.face_area
    ni, nc = m.n_interior, m.n_cells
    prim, grad = arr["prim"].reshape(5, nc), arr["grad"].reshape(15, nc)
    phi = arr["phi"].reshape(5, nc)
    bc = np.asarray(c.bc, dtype=np.float64).reshape(5, m.n_boundary)

    def side(cell, r):
        w = prim[:, cell]
        return reconstruct(w, grad[:, cell], phi[:, cell], r, pref) if second else w

    checked = total = 0
    for f in range(ni):
        o, nb = m.owner[f], m.neighbour[f]
        wl, wr = side(o, rl[:, f]), side(nb, rr[:, f])
        n = nrm[:, f]
        fl = lambda w: area[f] * orc.flux(kind, w, wr, n, sc.gamma, sc.mach_ref, pref)  # noqa: E731
        fr = lambda w: area[f] * orc.flux(kind, wl, w, n, sc.gamma, sc.mach_ref, pref)  # noqa: E731
        checked += check_block(parts["face_jac"][0, :, :, f], fl, wl)
        checked += check_block(parts["face_jac"][1, :, :, f], fr, wr)
        total += 10
    kinds = np.asarray(c.bc_kind, dtype=np.int64)
    for b in range(m.n_boundary):
        f = ni + b
        wf = side(m.owner[f], rl[:, f])
        n = nrm[:, f]
        got = parts["bnd_jac"][:, :, b]
        if kinds[b] in (meshes.SLIP, meshes.WALL):  # the pressure flux alone
            want = np.zeros((5, 5))
            want[1:4, 4] = area[f] * n
            np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
            continue
        data = bc[:, b]

        def fb(w, data=data, n=n, k=kinds[b], a=area[f]):
            wb = orc.boundary_value(w, k, data, n, sc.gamma, sc.gas_constant, pref)
            return a * orc.flux(kind, w, wb, n, sc.gamma, sc.mach_ref, pref)

        checked += check_block(got, fb, wf)
        total += 5
    assert checked >= 0.97 * total, (checked, total)


def test_dwdu_is_the_primitives_derivative():
    c = cases.tets_case(3)
    m, sc = c.mesh, c.scheme
    dev = DeviceStepper(m, sc, "fp64", c.bc_kind, c.bc, c.source)
    parts = dev.residual_and_jacobian(c.u)
    dev.close()
    u = np.asarray(c.u).reshape(5, -1)
    ref = np.asarray(sc.reference)
    for cell in range(0, m.n_cells, 17):
        q = u[:, cell] - ref

        def prim(qq):
            return orc.to_primitive(qq, ref, sc.gamma)

        check_block(parts["dwdu"][:, :, cell], prim, q, tol=1e-7)


@pytest.mark.parametrize("flux", ["hllc", "roe", "ausm+up"])
def test_first_order_matrix_is_the_residuals_jacobian(flux):
    """First order, every boundary a `state` face: A V = (R(U + e V) - R(U - e V)) / 2e."""
    c = cases.tets_case(4, flux=flux, second_order=False)
    m, sc = c.mesh, c.scheme
    print(f"jacobian matrix {flux}", flush=True)
    dev = DeviceStepper(m, sc, "fp64", c.bc_kind, c.bc, c.source)
    parts = dev.residual_and_jacobian(c.u)
    a = jacobian.assemble(m, parts)
    n = m.n_cells
    rng = np.random.default_rng(2)
    v = rng.normal(size=(5, n)) * np.array([1.0, 0.3, 0.3, 0.3, 2.5])[:, None]
    u = np.asarray(c.u).reshape(5, n)

    def central(e):
        return (dev.residual(u + e * v) - dev.residual(u - e * v)).reshape(5, n) / (2 * e)

    d, d4 = central(1e-6), central(0.25e-6)
    dev.close()
    av = jacobian.to_components(a @ jacobian.to_cells(v, n), n)
    scale = np.abs(d4).max(axis=1, keepdims=True)
    # A cell whose two differences disagree has a face state within the step of a switch of the
    # flux (HLLC's wave-speed min / max): a kink, not checked.
    smooth = np.all(np.abs(d - d4) <= 1e-6 * scale, axis=0)
    assert smooth.mean() >= 0.99, smooth.mean()
    np.testing.assert_allclose((av / scale)[:, smooth], (d4 / scale)[:, smooth], rtol=0, atol=2e-6)


def test_fp32_parts_are_the_fp64_parts_to_fp32_precision():
    c = cases.tets_case(4)
    out = {}
    for policy in ("fp64", "fp32+acc64"):
        dev = DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc, c.source)
        out[policy] = dev.residual_and_jacobian(c.u)
        dev.close()
    for k in ("dwdu", "face_jac", "bnd_jac"):
        a, b = out["fp64"][k], out["fp32+acc64"][k]
        assert np.abs(a - b).max() <= 1e-4 * np.abs(a).max(), k


def test_warped_meshes_carry_the_warp_terms_derivative():
    """On the warped polyhedral fixture the face blocks hold the warp term's derivative: they
    differ from the flux's alone, which the same faces give once their moments are zero."""
    c = cases.fixture_case("poly", "hllc")
    print("warped jacobian poly hllc fp64", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, "fp64", c.bc_kind, c.bc, c.source)
    warped = dev.residual_and_jacobian(c.u)["face_jac"]
    dev.close()
    flat = c.mesh
    flat.arrays["face_moment"] = np.zeros_like(flat.arrays["face_moment"])
    dev = DeviceStepper(flat, c.scheme, "fp64", c.bc_kind, c.bc, c.source)
    plain = dev.residual_and_jacobian(c.u)["face_jac"]
    dev.close()
    assert np.all(np.isfinite(warped))
    assert np.abs(warped - plain).max() > 1e-6 * np.abs(plain).max()


def test_the_implicit_solve_converges_in_few_solves():
    """First order HLLC on tets with sources and `state` boundaries (the Jacobian exact there):
    pseudo-transient continuation on the device's residual and Jacobian drops every residual 8
    orders within 60 linear solves."""
    from autan_amd import implicit

    c = cases.tets_case(4, flux="hllc", second_order=False)
    print("implicit solve tets 4 hllc first order", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, "fp64", c.bc_kind, c.bc, c.source)
    out = implicit.solve(dev, c.u, max_iterations=60, residual_orders=8.0, cfl0=10.0)
    dev.close()
    print(f"{out.iterations} solves, residual drop {np.log10(out.residuals[0] / out.residuals[-1])}", flush=True)
    assert out.converged, out.residuals[-1] / out.residuals[0]
