"""The steady implicit path's pieces on the CPU oracle, each against what the mathematics fixes:

- the warped-face Jacobian blocks (the face body's warp term differentiated through its state
  (W_L + W_R) / 2, kernels/euler.py:704-705 within 858-860) against central differences of the
  face body evaluated from the oracle's own flux and warp bodies, gradients and limiter held;
- the sweeps' colouring: a distance-1 colouring, each cell the least colour its lower-numbered face
  neighbours left (BlockMatrix's greedy colouring in cell order);
- the solve: the linear residual it reports is the one its update leaves in the assembled matrix
  (scipy), and it reaches GCR's 0.1 unless it used all 10 directions;
- the admissible step: the largest fraction of du keeping density and pressure above half, by its
  defining properties;
- the CFL control on prescribed residual sequences (switched evolution relaxation with its
  penalty), through the library's driver over a scripted target;
- a whole run on the oracle converging to 1e-8 on a sweep channel."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import cases
from autan_amd import jacobian, steady, sweep
from autan_amd import mesh as meshes
from autan_amd.stepper import FLUXES
from oracle import oracle as orc


def p_ref(sc) -> float:
    r = sc.reference
    return (sc.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0])


def reconstruct(w, g, lim, r, pref):
    d = np.array([g[3 * k] * r[0] + g[3 * k + 1] * r[1] + g[3 * k + 2] * r[2] for k in range(5)])
    f = w + lim * d
    return f if (f[0] > 0.0 and f[4] + pref > 0.0) else w


STATE_ZONES = [meshes.STATE] * 6


@pytest.mark.parametrize("flux", ["roe", "hllc", "ausm+up"])
def test_warped_face_blocks_are_central_differences_of_the_face_body(flux):
    c = cases.fixture_case("poly", flux=flux, kinds_by_zone=STATE_ZONES)
    m, sc = c.mesh, c.scheme
    assert not m.all_planar
    ora = orc.Oracle(m, sc, "fp64", c.bc_kind, c.bc, c.source)
    parts = ora.residual_and_jacobian(c.u)
    arr = ora.arrays()
    n, ni = m.n_cells, m.n_interior
    w, g, phi = arr["prim"].reshape(5, n), arr["grad"].reshape(15, n), arr["phi"].reshape(5, n)
    mom = np.asarray(m.face_moment).reshape(9, -1)
    nrm, area = np.asarray(m.face_normal).reshape(3, -1), np.asarray(m.face_area)
    rl, rr = np.asarray(m.r_left).reshape(3, -1), np.asarray(m.r_right).reshape(3, -1)
    pr, kind = p_ref(sc), FLUXES[flux]
    fj = parts["face_jac"].reshape(2, 5, 5, ni)
    warped = np.flatnonzero(np.abs(mom[:, :ni]).max(axis=0) > 0)
    rng = np.random.default_rng(3)
    moved = 0
    for f in rng.choice(warped, size=12, replace=False):
        lo, ro = m.owner[f], m.neighbour[f]

        def body(wl, wr):
            fl = reconstruct(wl, g[:, lo], phi[:, lo], rl[:, f], pr)
            fr = reconstruct(wr, g[:, ro], phi[:, ro], rr[:, f], pr)
            out = area[f] * orc.flux(kind, fl, fr, nrm[:, f], sc.gamma, sc.mach_ref, pr)
            return out + orc.warp(0.5 * (fl + fr), g[:, lo], phi[:, lo], g[:, ro], phi[:, ro], mom[:, f], sc.gamma, pr)

        for side in (0, 1):
            want = np.empty((5, 5))
            for j in range(5):
                e = np.zeros(5)
                h = 1e-6 * max(1.0, abs((w[:, lo] if side == 0 else w[:, ro])[j]))
                e[j] = h
                a = (w[:, lo] + e, w[:, ro]) if side == 0 else (w[:, lo], w[:, ro] + e)
                b = (w[:, lo] - e, w[:, ro]) if side == 0 else (w[:, lo], w[:, ro] - e)
                want[:, j] = (body(*a) - body(*b)) / (2 * h)
            got = fj[side, :, :, f]
            scale = np.abs(want).max()
            assert np.abs(got - want).max() <= 2e-7 * scale, (f, side)
            # The warp term's own derivative is in the block: the flux's alone differs from it.
            fl = reconstruct(w[:, lo], g[:, lo], phi[:, lo], rl[:, f], pr)
            fr = reconstruct(w[:, ro], g[:, ro], phi[:, ro], rr[:, f], pr)
            plain = area[f] * orc.face_jacobian(kind, fl, fr, nrm[:, f], sc.gamma, sc.mach_ref, pr)[side]
            moved += int(np.abs(plain - got).max() > 1e-9 * scale)
    assert moved > 0


def test_the_sweeps_colouring_is_the_greedy_distance_one_colouring():
    c = cases.fixture_case("poly", kinds_by_zone=STATE_ZONES)
    m = c.mesh
    ora = orc.Oracle(m, c.scheme, "fp64", c.bc_kind, c.bc, c.source)
    parts = ora.residual_and_jacobian(c.u)
    _, _, _, col = orc.implicit_solve(m, parts, ora.spectral_radius(), 100.0, c.scheme.gamma, c.u)
    ni = m.n_interior
    own, nei = np.asarray(m.owner[:ni]), np.asarray(m.neighbour)
    assert np.all(col[own] != col[nei])
    adj = [set() for _ in range(m.n_cells)]
    for a, b in zip(own.tolist(), nei.tolist()):
        adj[a].add(b)
        adj[b].add(a)
    for cell in range(m.n_cells):
        taken = {int(col[o]) for o in adj[cell] if o < cell}
        assert int(col[cell]) == min(set(range(len(taken) + 1)) - taken)


@pytest.mark.parametrize("cfl", [10.0, 1e4])
@pytest.mark.parametrize("second", [False, True])
def test_the_solve_reports_the_linear_residual_its_update_leaves(cfl, second):
    c = cases.fixture_case("poly", flux="hllc", second_order=second, kinds_by_zone=STATE_ZONES)
    m, n = c.mesh, c.mesh.n_cells
    ora = orc.Oracle(m, c.scheme, "fp64", c.bc_kind, c.bc, c.source)
    parts = ora.residual_and_jacobian(c.u)
    lam = ora.spectral_radius()
    u1, du, rep, _ = orc.implicit_solve(m, parts, lam, cfl, c.scheme.gamma, c.u)
    a = jacobian.assemble(m, {"dwdu": parts["dwdu"].reshape(5, 5, n),
                              "face_jac": parts["face_jac"].reshape(2, 5, 5, -1),
                              "bnd_jac": parts["bnd_jac"].reshape(5, 5, -1)}, shift=lam / cfl)
    b = -jacobian.to_cells(parts["res"], n)
    left = a @ jacobian.to_cells(du, n) - b
    drop = np.linalg.norm(left) / np.linalg.norm(b)
    assert rep["residual_drop"] == pytest.approx(drop, rel=1e-8)
    assert rep["residual_drop"] <= 0.1 or rep["cycles"] == 10
    assert not rep["failed"] and 0.0 < rep["alpha"] <= 1.0
    np.testing.assert_allclose(u1, np.asarray(c.u) + rep["alpha"] * du, rtol=0, atol=1e-15 * np.abs(c.u).max())


def test_the_admissible_step_is_the_largest_fraction_keeping_half_the_density_and_pressure():
    g = 1.4
    rng = np.random.default_rng(8)
    n = 400
    rho = rng.uniform(0.5, 2.0, n)
    v = rng.normal(0.0, 0.3, (3, n))
    p = rng.uniform(0.5, 1.5, n)
    u = np.vstack([rho, rho * v, p / (g - 1.0) + 0.5 * rho * (v * v).sum(axis=0)])

    def pressure(x):
        return (g - 1.0) * (x[4] - 0.5 * (x[1] ** 2 + x[2] ** 2 + x[3] ** 2) / x[0])

    # No decrease: the whole step.
    assert orc.admissible_step(u, 0.1 * np.abs(u), g) == 1.0
    for trial in range(20):
        du = rng.normal(0.0, 1.0, (5, n)) * np.abs(u) * rng.uniform(0.2, 3.0)
        alpha = orc.admissible_step(u, du, g)
        nxt = u + alpha * du
        assert np.all(nxt[0] >= 0.5 * rho * (1 - 1e-14))
        assert np.all(pressure(nxt) >= 0.5 * p * (1 - 1e-14))
        if alpha == 1.0:
            continue
        # The limiting cell: its density lands on half its value, or twice the fraction breaks its
        # pressure bound (the fraction halves from the density bound until the pressure holds).
        a_rho = np.where(du[0] < 0, 0.5 * rho / np.where(du[0] < 0, -du[0], 1.0), np.inf)
        hit_rho = np.isclose(alpha, a_rho, rtol=1e-14)
        twice = u + 2 * alpha * du
        hit_p = ~(pressure(twice) >= 0.5 * p) & (2 * alpha <= np.minimum(1.0, a_rho) * (1 + 1e-14))
        assert np.any(hit_rho | hit_p), trial


class Scripted:
    """A target whose residuals follow a script and whose solves report what the script says."""

    def __init__(self, norms, reports):
        self.norms, self.reports, self.i = norms, reports, 0
        self.cfls = []

    def evaluate(self):
        return self.norms[min(self.i, len(self.norms) - 1)], None

    def set_second_order(self, on):
        pass

    def jacobian(self):
        return self.norms[min(self.i, len(self.norms) - 1)], None

    def solve(self, cfl):
        self.cfls.append(cfl)
        rep = self.reports[self.i]
        self.i += 1
        return rep


def test_the_cfl_number_follows_the_switched_evolution_relaxation():
    # Second order from the first step, one equation set: drop = r / r_0 per iteration.
    r0 = np.ones(5)
    drops = [1.0, 0.1, 0.01, 0.02, 0.015, 0.001, 0.0005, 0.0004, 0.0001]
    ok = [3, 0.05, 1.0, 0.0, 0.0]
    short = [3, 0.05, 0.25, 0.0, 0.0]
    fail = [10, 0.95, 0.0, 1.0, 0.0]
    reports = [ok, ok, ok, ok, short, ok, fail, ok, ok]
    t = Scripted([r0 * d for d in drops], reports)
    res = steady.solve_implicit_with(t, max_iterations=len(drops), residual_orders=20.0, start_order="second")
    # Step by step (penalty 1 at the start; 1.5x recovery capped at 1; halved after a rise above
    # 1.1x the previous drop or after a shortened or rejected step; CFL = penalty 1000 / drop,
    # capped after a shortened or rejected step at half the CFL used, clamped to [1, 1e6]).
    want = [
        1000.0,  # drop 1, penalty 1
        1e4,  # drop 0.1
        1e5,  # drop 0.01
        0.5 * 1000.0 / 0.02,  # drop 0.02 rose: penalty 0.5 -> 25000
        0.75 * 1000.0 / 0.015,  # penalty 0.75 -> 50000; this step is shortened (alpha 0.25)
        # shortened: penalty = min(0.75, 50000 * 0.015 / 1000) = 0.75, cap 25000; now halved: 0.375
        min(0.375 * 1000.0 / 0.001, 25000.0),
        # penalty recovers to 0.5625: 1125000, clamped to 1e6; this step fails: penalty
        # min(0.5625, 1e6 * 0.0005 / 1000) = 0.5, cap 500000
        1e6,
        min(0.25 * 1000.0 / 0.0004, 500000.0),  # halved after the rejected step: 0.25
        min(0.375 * 1000.0 / 0.0001, 1e6),  # recovered by half: 0.375, clamped
    ]
    assert res.iterations == len(drops)
    np.testing.assert_allclose(t.cfls, want, rtol=1e-14)
    np.testing.assert_allclose(res.cfl, want, rtol=1e-14)


def test_an_oracle_implicit_run_converges_on_a_sweep_channel():
    orc.set_threads(8)
    try:
        sp = dataclasses.replace(sweep.specs(64, seed=1)[0], lattice=sweep.lattice_for(1500))
        c = sweep.build(sp, "fp64")
        ora = orc.Oracle(c.mesh, c.scheme, "fp64", c.bc_kind, c.bc)
        t = cases.OracleImplicitTarget(ora, c.u, c.mesh, c.scheme.gamma)
        r = steady.solve_implicit_with(t, max_iterations=150, residual_orders=8.0)
    finally:
        orc.set_threads(1)
    assert r.status == steady.CONVERGED and r.iterations < 100
    assert r.first_order_end == "drop"
    assert r.cfl[0] == 1000.0 and r.cfl.max() == 1e6


@pytest.mark.parametrize("flux", ["hllc", "ausm+up"])
def test_the_boundary_gradient_term_is_the_boundary_fluxes_derivative_through_the_gradients(flux):
    """B v (the Krylov operator's boundary term) against central differences of the boundary faces'
    fluxes in which only the gradients move: each boundary-owning cell's gradient refitted from
    W + e dW (dW = dW/dq v, boundary values held), its primitives and limiter held."""
    zones = [meshes.FARFIELD, meshes.OUTLET, meshes.SLIP, meshes.WALL, meshes.INLET, meshes.STATE]
    c = cases.fixture_case("poly", flux=flux, kinds_by_zone=zones)
    m, sc = c.mesh, c.scheme
    n, ni, nb = m.n_cells, m.n_interior, m.n_boundary
    ora = orc.Oracle(m, sc, "fp64", c.bc_kind, c.bc, c.source)
    parts = ora.residual_and_jacobian(c.u)
    arr = ora.arrays()
    w, g, phi = arr["prim"].reshape(5, n), arr["grad"].reshape(15, n), arr["phi"].reshape(5, n)
    v = np.random.default_rng(5).normal(size=(5, n)) * np.array([1.0, 0.3, 0.3, 0.3, 2.0])[:, None]
    # B v: the term alone, through a solve-free call (y starts at zero).
    got = np.zeros(5 * n)
    lib = orc.lib()
    import ctypes
    lib.oracle_boundary_term.argtypes = [ctypes.c_void_p] * 5
    vv = np.ascontiguousarray(v.reshape(-1))
    lib.oracle_boundary_term(ora.h, vv.ctypes.data, parts["dwdu"].ctypes.data, parts["bnd_jac"].ctypes.data,
                             got.ctypes.data)
    got = got.reshape(5, n)
    dwdu = parts["dwdu"].reshape(5, 5, n)
    dw = np.einsum("ijc,jc->ic", dwdu, v)
    so, st = np.asarray(m.stencil_offsets), np.asarray(m.stencil)
    sw = np.asarray(m.stencil_weights).reshape(3, -1)
    nrm, area = np.asarray(m.face_normal).reshape(3, -1), np.asarray(m.face_area)
    rl, mom = np.asarray(m.r_left).reshape(3, -1), np.asarray(m.face_moment).reshape(9, -1)
    pr, kind = p_ref(sc), FLUXES[flux]
    bc = np.asarray(c.bc).reshape(5, nb)

    def grad_of(cell, eps):
        wc = w[:, cell] + eps * dw[:, cell]
        out = np.zeros(15)
        for k in range(so[cell], so[cell + 1]):
            e = st[k]
            # a boundary entry -(f + 1) holds face f's boundary state
            wn = (w[:, e] + eps * dw[:, e]) if e >= 0 else arr["prim_b"].reshape(5, nb)[:, -e - 1 - ni]
            for i in range(5):
                out[3 * i : 3 * i + 3] += (wn[i] - wc[i]) * sw[:, k]
        return out

    def boundary_fluxes(cell, gr):
        out = np.zeros(5)
        for k in range(m.cell_face_offsets[cell], m.cell_face_offsets[cell + 1]):
            f = m.cell_faces[k]
            if f < ni:
                continue
            b = f - ni
            wf = reconstruct(w[:, cell], gr, phi[:, cell], rl[:, f], pr)
            kd = c.bc_kind[b]
            if kd in (meshes.SLIP, meshes.WALL):
                fx = area[f] * np.array([0.0, *(wf[4] * nrm[:, f]), 0.0])
                M = mom[:, f].reshape(3, 3)
                fx[1:4] += phi[4, cell] * (M.T @ gr[12:15])
            else:
                wb = orc.boundary_value(wf, kd, bc[:, b], nrm[:, f], sc.gamma, sc.gas_constant, pr)
                fx = area[f] * orc.flux(kind, wf, wb, nrm[:, f], sc.gamma, sc.mach_ref, pr)
            out += fx
        return out

    owners = sorted(set(np.asarray(m.owner[ni:]).tolist()))
    checked = 0
    for cell in owners[::3]:
        h = 1e-6
        fd = (boundary_fluxes(cell, grad_of(cell, h)) - boundary_fluxes(cell, grad_of(cell, -h))) / (2 * h)
        scale = np.abs(fd).max() + 1e-12
        assert np.abs(got[:, cell] - fd).max() <= 1e-6 * scale, cell
        checked += int(np.abs(fd).max() > 1e-8)
    assert checked > 5
    # Cells that own no boundary face get nothing.
    inner = np.setdiff1d(np.arange(n), owners)
    assert np.all(got[:, inner] == 0.0)
