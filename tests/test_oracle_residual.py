"""The oracle's residual, spectral radius, step and norms against properties the reference's own
host tests hold (tests/physics/test_euler.cpp), on its test meshes (tests/data/mesh, copied under
tests/golden/meshes): each runs in fp64 and fp32+acc64 with Roe, HLLC and AUSM+-up.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd.stepper import Scheme
from oracle import oracle as orc
from oracle.oracle import Oracle

GAS_R = 287.0
MESHES = ["hex", "tet-gmsh", "poly", "hybrid"]
POLICIES = ["fp64", "fp32+acc64"]
FLUXES = ["roe", "hllc", "ausm+up"]
EPS = {"fp64": np.finfo(np.float64).eps, "fp32+acc64": np.finfo(np.float32).eps}


def conserved(rho, v, p, gamma=1.4):
    v = np.asarray(v, dtype=float)
    return np.array([rho, *(rho * v), p / (gamma - 1.0) + 0.5 * rho * (v @ v)])


def scheme(flux, mach, limiter=True):
    """test_euler.cpp:140-153: a reference state apart from every test state."""
    return Scheme(
        flux=flux,
        second_order=True,
        limiter=limiter,
        gamma=1.4,
        gas_constant=GAS_R,
        mach_ref=mach,
        scale=(1.0, 100.0, 100.0, 100.0, 1e5),
        reference=tuple(conserved(1.1, (10.0, 5.0, 0.0), 0.95e5)),
    )


_MESHES: dict = {}


def mesh(name):
    if name not in _MESHES:
        _MESHES[name] = meshes.read_polymesh(cases.GOLDEN / "meshes" / name)
    return _MESHES[name]


def surface(m):
    s = np.zeros(m.n_cells)
    np.add.at(s, m.owner, m.face_area)
    np.add.at(s, m.neighbour, m.face_area[: m.n_interior])
    return s


def relative_residual(r, m, rho, v, p, gamma=1.4):
    """test_euler.cpp:104-127: max |R_k| over the flux bound of the uniform state times the
    cell's surface."""
    speed = float(np.linalg.norm(v)) + math.sqrt(gamma * p / rho)
    uc = conserved(rho, v, p)
    flux = np.array([rho * speed, rho * speed**2 + p, rho * speed**2 + p, rho * speed**2 + p, (uc[4] + p) * speed])
    rel = np.abs(r.reshape(5, -1)) / (flux[:, None] * surface(m)[None, :])
    return float(rel.max()) if np.all(np.isfinite(rel)) else math.inf


def uniform(m, rho, v, p):
    return np.repeat(conserved(rho, v, p)[:, None], m.n_cells, axis=1).reshape(-1)


def zone_names(m):
    names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    return [names[z] for z in m.face_zone]


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("name", MESHES)
def test_uniform_flow_inlet_outlet_slip_is_steady(name, flux, policy):
    """test_euler.cpp:158-187: total conditions in, static pressure out, slip walls along the flow."""
    m = mesh(name)
    rho, p = 1.2, 1e5
    a = math.sqrt(1.4 * p / rho)
    v = np.array([0.5 * a, 0.0, 0.0])
    t = p / (rho * GAS_R)
    t0 = t * (1.0 + 0.2 * 0.25)
    p0 = p * (t0 / t) ** 3.5
    z = zone_names(m)
    kinds = np.array([meshes.INLET if s == "xmin" else meshes.OUTLET if s == "xmax" else meshes.SLIP for s in z])
    data = np.zeros((5, m.n_boundary))
    data[0, kinds == meshes.INLET] = p0
    data[1, kinds == meshes.INLET] = t0
    data[4, kinds == meshes.OUTLET] = p
    meshes.set_boundary_kinds(m, kinds)
    o = Oracle(m, scheme(flux, 0.5), policy, kinds.astype(float), data)
    r = o.residual(uniform(m, rho, v, p))
    assert relative_residual(r, m, rho, v, p) <= 64 * EPS[policy]


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("mach", [0.5, 2.0])
@pytest.mark.parametrize("name", MESHES)
def test_uniform_oblique_flow_with_farfield_is_steady(name, mach, flux, policy):
    """test_euler.cpp:189-209: farfield on every face, subsonic and supersonic, oblique flow."""
    m = mesh(name)
    rho, p = 1.2, 1e5
    a = math.sqrt(1.4 * p / rho)
    v = np.array([mach * a * 0.8, mach * a * 0.36, mach * a * 0.48])
    kinds = np.full(m.n_boundary, meshes.FARFIELD)
    data = np.repeat(np.array([rho, *v, p])[:, None], m.n_boundary, axis=1)
    meshes.set_boundary_kinds(m, kinds)
    o = Oracle(m, scheme(flux, mach), policy, kinds.astype(float), data)
    r = o.residual(uniform(m, rho, v, p))
    assert relative_residual(r, m, rho, v, p) <= 64 * EPS[policy]


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("name", MESHES)
def test_gas_at_rest_with_a_density_jump_is_steady(name, flux, policy):
    """test_euler.cpp:212-240: a stationary contact (10:1 density jump) under slip walls."""
    m = mesh(name)
    kinds = np.full(m.n_boundary, meshes.SLIP)
    meshes.set_boundary_kinds(m, kinds)
    x = m.centroids()
    rho = 1.0 + 0.5 * np.sin(7.0 * x[:, 0] + 3.0 * x[:, 1]) * np.cos(5.0 * x[:, 2]) + np.where(x[:, 0] > 0.5, 9.0, 0.0)
    p = 1e5
    u = np.stack([rho, 0 * rho, 0 * rho, 0 * rho, np.full_like(rho, p / 0.4)]).reshape(-1)
    o = Oracle(m, scheme(flux, 0.1), policy, kinds.astype(float), np.zeros((5, m.n_boundary)))
    r = o.residual(u)
    assert relative_residual(r, m, 11.0, np.zeros(3), p) <= 64 * EPS[policy]


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("name", MESHES)
def test_linear_velocity_has_the_exact_mass_divergence(name, policy):
    """test_euler.cpp:430-469: rho = 1, v = (a x, 0, 0) tangent to slip sides, exact state on the
    x sides, unlimited: the linear reconstruction is exact and R_mass = a V_c."""
    m = mesh(name)
    a, p = 50.0, 1e5
    z = zone_names(m)
    kinds = np.array([meshes.STATE if s in ("xmin", "xmax") else meshes.SLIP for s in z])
    xf = m.face_centroids()[m.n_interior :]
    data = np.stack([np.ones(m.n_boundary), a * xf[:, 0], 0 * xf[:, 0], 0 * xf[:, 0], np.full(m.n_boundary, p)])
    meshes.set_boundary_kinds(m, kinds)
    x = m.centroids()
    u = np.stack([np.ones(m.n_cells), a * x[:, 0], 0 * x[:, 0], 0 * x[:, 0], p / 0.4 + 0.5 * (a * x[:, 0]) ** 2]).reshape(-1)
    o = Oracle(m, scheme("roe", 0.3, limiter=False), policy, kinds.astype(float), data)
    r = o.residual(u).reshape(5, -1)
    assert np.all(np.abs(r[0] - a * m.cell_volume) <= 64 * EPS[policy] * a * surface(m))


@pytest.mark.parametrize("name", MESHES)
def test_spectral_radius_closed_form(name):
    """test_euler.cpp:471-518: opposing flows; sum over faces of max over both sides of
    (|v . n| + c) A, slip walls mirroring the cell."""
    m = mesh(name)
    kinds = np.full(m.n_boundary, meshes.SLIP)
    meshes.set_boundary_kinds(m, kinds)
    n = m.n_cells
    s = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    rho = 1.0 + 0.1 * (np.arange(n) % 3)
    v = np.stack([1000.0 * s, 300.0 * s, -200.0 * s], axis=1)
    p = 1e5
    u = np.stack([rho, rho * v[:, 0], rho * v[:, 1], rho * v[:, 2], p / 0.4 + 0.5 * rho * (v * v).sum(1)]).reshape(-1)
    o = Oracle(m, scheme("roe", 0.5), "fp64", kinds.astype(float), np.zeros((5, m.n_boundary)))
    o.residual(u)
    lam = o.spectral_radius()
    nrm = m.face_normal.reshape(3, -1).T
    area = m.face_area
    c = np.sqrt(1.4 * p / rho)

    def speed(i, f):
        return (abs(v[i] @ nrm[f]) + c[i]) * area[f]

    expect = np.zeros(n)
    for f in range(m.n_interior):
        o_, n_ = m.owner[f], m.neighbour[f]
        mx = max(speed(o_, f), speed(n_, f))
        expect[o_] += mx
        expect[n_] += mx
    for f in range(m.n_interior, m.n_faces):
        expect[m.owner[f]] += speed(m.owner[f], f)
    np.testing.assert_allclose(lam, expect, rtol=1e-12)


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("name", MESHES)
def test_least_squares_gradient_is_exact_on_linear_fields(name, policy):
    """A linear primitive field with its exact values on `state` faces: every cell's gradient is
    the field's (the least-squares fit is exact on linear data), its bounds bracket the cell."""
    m = mesh(name)
    kinds = np.full(m.n_boundary, meshes.STATE)
    meshes.set_boundary_kinds(m, kinds)
    grad = np.array([[0.3, -0.2, 0.1], [2.0, 1.0, -1.0], [-1.5, 0.5, 2.5], [0.7, -0.3, 1.1], [3000.0, -2000.0, 1000.0]])
    w0 = np.array([1.2, 10.0, 5.0, 1.0, 1e5])

    def field(x):
        return (w0[:, None] + grad @ x.T)

    xf = m.face_centroids()[m.n_interior :]
    w = field(m.centroids())
    u = np.stack([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3], w[4] / 0.4 + 0.5 * w[0] * (w[1:4] ** 2).sum(0)]).reshape(-1)
    o = Oracle(m, scheme("hllc", 0.3, limiter=True), "fp64" if policy == "fp64" else policy, kinds.astype(float), field(xf))
    o.residual(u)
    a = o.arrays()
    g = a["grad"].reshape(15, -1)
    prim = a["prim"].reshape(5, -1)
    # The stored primitives of the state (u is rounded to the policy's state first) carry
    # round-off of eps |W|; the fit turns that into eps |W| / h of the gradient.
    h = np.cbrt(m.cell_volume).min()
    for i in range(5):
        for j in range(3):
            tol = 400 * EPS[policy] * (abs(grad[i, j]) + np.abs(w[i]).max() / h)
            assert np.abs(g[3 * i + j] - grad[i, j]).max() <= tol, (i, j)
    assert np.all(a["wmin"].reshape(5, -1) <= prim) and np.all(prim <= a["wmax"].reshape(5, -1))


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("name", ["poly", "tet-gmsh"])
def test_gather_and_scatter_orders_agree(name, flux, policy):
    """The store-gather order of the device path against the host's face-order scatter: the
    reference's 500 eps f_k A_i unit (tests/physics/test_euler_device.cpp:68-211)."""
    case = cases.fixture_case(name, flux)
    o = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    a = o.residual(case.u, gather=True)
    b = o.residual(case.u, gather=False)
    units = cases.normwise_units(case, EPS[policy]).reshape(-1)
    assert np.max(np.abs(a - b) / units) <= 500


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("name", MESHES)
def test_discrete_conservation(name, policy):
    """Sum over the cells of the face part of R equals the sum of the boundary fluxes: interior
    fluxes cancel pairwise (the conservation check of docs/design/euler.md)."""
    case = cases.fixture_case(name, "hllc")
    o = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, None)
    r = o.residual(case.u).reshape(5, -1)
    f = o.face_fluxes().reshape(5, -1)
    m = case.mesh
    total = r.sum(axis=1)
    boundary = f[:, m.n_interior :].sum(axis=1)
    scale = np.abs(f).sum(axis=1)
    assert np.all(np.abs(total - boundary) <= 1e-13 * scale), (total, boundary)
    # and the residual is that face part: every interior flux once with each sign
    recon = np.zeros_like(r)
    for k in range(5):
        np.add.at(recon[k], m.owner, f[k])
        np.subtract.at(recon[k], m.neighbour, f[k, : m.n_interior])
    assert np.all(np.abs(recon - r) <= 1e-12 * scale[:, None])


@pytest.mark.parametrize("policy", POLICIES)
def test_ssp_rk3_stages_integrate_a_constant_source_exactly(policy):
    """With a uniform flow the face part of R vanishes on a periodic hexahedral mesh, R = -V s is
    constant while the state stays uniform, and the three stages (0, 1), (3/4, 1/4), (1/3, 2/3) of Shu and Osher (1988) advance u
    by exactly dt s; norms are sqrt(sum (V s)^2 / N)."""
    m = meshes.periodic_hex(6, 0.0, 1.0)
    sc = scheme("roe", 0.5)
    sc.second_order = False
    sc.limiter = False
    # One source for every cell, so that the state stays uniform through the stages.
    s = np.repeat(np.array([1e-3, -2e-3, 3e-3, 1e-3, 0.5])[:, None], m.n_cells, axis=1)
    u = uniform(m, 1.1, (10.0, 5.0, 0.0), 0.95e5)
    o = Oracle(m, sc, policy, np.zeros(0), np.zeros((5, 0)), s)
    r = o.residual(u).reshape(5, -1)
    # The face part is round-off of the face fluxes it sums (relative_residual's unit).
    face = r + m.cell_volume[None, :] * (s if policy == "fp64" else s.astype(np.float32))
    assert relative_residual(face, m, 1.1, np.array([10.0, 5.0, 0.0]), 0.95e5) <= 64 * EPS[policy]
    dt = 1e-3
    new, norms, h, inv = o.step(u, dt=dt)
    assert h == dt and inv is None
    du = (new - u).reshape(5, -1)
    # Three stages of a constant R: u + dt (-R / V), exactly but for the last bits of u.
    np.testing.assert_allclose(du, -dt * r / m.cell_volume[None, :], rtol=0, atol=4 * np.finfo(float).eps * np.abs(u).max())
    expect = np.sqrt((r**2).sum(axis=1) / m.n_cells)
    np.testing.assert_allclose(norms, expect, rtol=1e-15)


def test_first_invalid_cell_is_the_least():
    """EulerPhysics::first_invalid_cell: a non-positive density or pressure as the kernels read
    the stored state, else a non-finite residual; the least cell."""
    case = cases.fixture_case("hex", "roe")
    o = Oracle(case.mesh, case.scheme, "fp32+acc64", case.bc_kind, case.bc)
    u = case.u.reshape(5, -1).copy()
    assert o.first_invalid(u) is None
    u[0, [9, 4, 20]] = -1.0
    assert o.first_invalid(u) == 4
    u = case.u.reshape(5, -1).copy()
    u[4, 7] = 0.0  # zero energy: negative pressure
    res = np.zeros_like(u)
    res[2, 3] = np.nan
    assert o.first_invalid(u, res) == 7
    assert o.first_invalid(case.u, res) == 3


def test_renumbering_moves_the_residual_with_the_cells():
    """The residual is a function of the mesh, not of its numbering: RCM and Morton orders give
    the same per-cell residual within the parity bound."""
    a = cases.tets_case(4, order="rcm")
    b = cases.tets_case(4, order="morton")
    key = lambda m: np.lexsort(np.round(m.centroids(), 12).T)  # noqa: E731
    ia, ib = key(a.mesh), key(b.mesh)
    np.testing.assert_allclose(a.mesh.centroids()[ia], b.mesh.centroids()[ib], rtol=0, atol=1e-14)
    ra = Oracle(a.mesh, a.scheme, "fp64", a.bc_kind, a.bc, a.source).residual(a.u).reshape(5, -1)[:, ia]
    rb = Oracle(b.mesh, b.scheme, "fp64", b.bc_kind, b.bc, b.source).residual(b.u).reshape(5, -1)[:, ib]
    units = cases.normwise_units(a, EPS["fp64"])[:, ia]
    assert np.max(np.abs(ra - rb) / units) <= 500


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_threaded_gather_loops_give_the_serial_bits(policy):
    """The gather variant on several threads keeps each cell's arithmetic and order, so the
    state, norms and time step are those of one thread, bit for bit."""
    case = cases.fixture_case("poly", "hllc")
    ora = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    out = []
    for t in (1, 4):
        orc.set_threads(t)
        try:
            u, norms, h, inv = ora.step(case.u.copy(), cfl=0.5)
            u, norms2, h2, inv = ora.step(u, cfl=0.5, local=True)
        finally:
            orc.set_threads(1)
        out.append((u, norms, h, norms2))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][3], out[1][3])
    assert out[0][2] == out[1][2]
