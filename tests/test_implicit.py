"""The implicit path on the CPU: a stepper made of the oracle's residual, its Jacobian blocks
(face_jacobian, boundary_jacobian, dwdu) and its spectral radius, at first order. The assembled
matrix is the residual's derivative (central differences of the oracle's residual; `state`
boundaries, whose values do not depend on the cells, make it exact), and the implicit solve
(autan_amd/implicit.py) converges on it as it does on the device."""

from __future__ import annotations

import numpy as np

import cases
from autan_amd import implicit, jacobian
from autan_amd.stepper import FLUXES
from oracle import oracle as orc


def p_ref(sc) -> float:
    r = sc.reference
    return (sc.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0])


def reconstruct(w, g, lim, r, pref):
    """kernels/euler.py:619-627: W + phi (grad W . r), the cell value where the density or pressure
    would not stay positive."""
    d = np.array([g[3 * k] * r[0] + g[3 * k + 1] * r[1] + g[3 * k + 2] * r[2] for k in range(5)])
    f = w + lim * d
    return f if (f[0] > 0.0 and f[4] + pref > 0.0) else w


class OracleStepper:
    """residual_and_jacobian and spectral_radius of the oracle in FP64: the Jacobian blocks at the
    reconstructed face states, the gradients and limiter held fixed."""

    def __init__(self, case, flux):
        self.case, self.mesh, self.kind = case, case.mesh, FLUXES[flux]
        self.second = case.scheme.second_order
        self.ora = orc.Oracle(case.mesh, case.scheme, "fp64", case.bc_kind, case.bc, case.source)

    def residual_and_jacobian(self, u):
        m, sc = self.mesh, self.case.scheme
        n, ni, nb = m.n_cells, m.n_interior, m.n_boundary
        ref = np.asarray(sc.reference)
        res = self.ora.residual(np.asarray(u).reshape(-1)).reshape(5, n)
        arr = self.ora.arrays()
        w, grad, phi = arr["prim"].reshape(5, n), arr["grad"].reshape(15, n), arr["phi"].reshape(5, n)
        q = np.asarray(u).reshape(5, n) - ref[:, None]
        dwdu = np.stack([orc.dwdu(q[:, c], ref, sc.gamma) for c in range(n)], axis=2)
        nrm, area, pr = m.face_normal.reshape(3, -1), m.face_area, p_ref(sc)
        rl, rr = m.r_left.reshape(3, -1), m.r_right.reshape(3, -1)

        def side(c, r):
            return reconstruct(w[:, c], grad[:, c], phi[:, c], r, pr) if self.second else w[:, c]

        fj = np.stack([area[f] * orc.face_jacobian(self.kind, side(m.owner[f], rl[:, f]), side(m.neighbour[f], rr[:, f]),
                                                   nrm[:, f], sc.gamma, sc.mach_ref, pr) for f in range(ni)], axis=3)
        bc = np.asarray(self.case.bc, dtype=np.float64).reshape(5, nb)
        bj = np.stack([area[ni + b] * orc.boundary_jacobian(self.kind, side(m.owner[ni + b], rl[:, ni + b]),
                                                            self.case.bc_kind[b], bc[:, b], nrm[:, ni + b], sc.gamma,
                                                            sc.gas_constant, sc.mach_ref, pr) for b in range(nb)], axis=2)
        return {"res": res, "dwdu": dwdu, "face_jac": fj, "bnd_jac": bj}

    def spectral_radius(self):
        return self.ora.spectral_radius()


def test_the_oracles_first_order_matrix_is_its_residuals_jacobian():
    c = cases.tets_case(3, flux="roe", second_order=False)
    st = OracleStepper(c, "roe")
    parts = st.residual_and_jacobian(c.u)
    a = jacobian.assemble(c.mesh, parts)
    n = c.mesh.n_cells
    v = np.random.default_rng(6).normal(size=(5, n)) * np.array([1.0, 0.3, 0.3, 0.3, 2.5])[:, None]
    u = np.asarray(c.u).reshape(5, n)

    def central(e):
        return (st.ora.residual((u + e * v).reshape(-1)) - st.ora.residual((u - e * v).reshape(-1))).reshape(5, n) / (2 * e)

    d, d4 = central(1e-6), central(0.25e-6)
    av = jacobian.to_components(a @ jacobian.to_cells(v, n), n)
    scale = np.abs(d4).max(axis=1, keepdims=True)
    smooth = np.all(np.abs(d - d4) <= 1e-6 * scale, axis=0)
    assert smooth.mean() >= 0.99
    np.testing.assert_allclose((av / scale)[:, smooth], (d4 / scale)[:, smooth], rtol=0, atol=2e-6)


def test_the_implicit_solve_converges_on_the_oracle():
    c = cases.tets_case(3, flux="hllc", second_order=False)
    out = implicit.solve(OracleStepper(c, "hllc"), c.u, max_iterations=40, residual_orders=8.0, cfl0=10.0)
    assert out.converged and out.iterations <= 20, (out.iterations, out.residuals[-1] / out.residuals[0])


def test_the_implicit_solve_converges_at_second_order_with_the_approximate_jacobian():
    """Second order: the Jacobian holds the reconstruction's gradients and limiter fixed, so the
    solve is no longer Newton's; it still drops 8 orders in a few dozen solves (25 here)."""
    c = cases.tets_case(3, flux="hllc", second_order=True)
    out = implicit.solve(OracleStepper(c, "hllc"), c.u, max_iterations=40, residual_orders=8.0, cfl0=10.0)
    assert out.converged, (out.iterations, out.residuals[-1] / out.residuals[0])


def test_gmres_with_block_jacobi_converges_as_the_exact_solves_do():
    """The linear systems solved to 1e-3 by GMRES preconditioned with the diagonal blocks' inverses:
    the outer iteration still drops 8 orders, in about as many solves as with exact LU."""
    c = cases.tets_case(3, flux="hllc", second_order=False)
    lu = implicit.solve(OracleStepper(c, "hllc"), c.u, max_iterations=40, residual_orders=8.0, cfl0=10.0)
    kr = implicit.solve(OracleStepper(c, "hllc"), c.u, max_iterations=40, residual_orders=8.0, cfl0=10.0,
                        linear="gmres")
    assert lu.converged and kr.converged, (kr.iterations, kr.residuals[-1] / kr.residuals[0])
    assert kr.iterations <= 2 * lu.iterations + 4, (kr.iterations, lu.iterations)


def test_gauss_seidel_sweeps_converge_as_the_reference_relaxes():
    """Four symmetric multicolour block Gauss-Seidel sweeps per system (the reference's relaxation,
    BlockMatrix::gauss_seidel) in place of the exact solve: still 8 orders, in more solves."""
    c = cases.tets_case(3, flux="hllc", second_order=False)
    out = implicit.solve(OracleStepper(c, "hllc"), c.u, max_iterations=60, residual_orders=8.0, cfl0=10.0,
                         linear="gs", sweeps=4)
    assert out.converged, (out.iterations, out.residuals[-1] / out.residuals[0])
