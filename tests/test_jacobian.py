"""Assembly of the residual's Jacobian (autan_amd/jacobian.py, Euler::assemble) on the host: the
block matrix applied to a vector equals the linearisation of the face loops written face by face
(R_owner += F(W_L, W_R), R_neighbour -= F, boundary R_owner += F(W_L)), with W = D U per cell."""

from __future__ import annotations

import numpy as np

from autan_amd import jacobian
from autan_amd import mesh as meshes


def test_the_block_matrix_is_the_face_loops_linearisation():
    m = meshes.kuhn_tets(2, 0.1, 0)
    n, ni, nb = m.n_cells, m.n_interior, m.n_boundary
    rng = np.random.default_rng(4)
    parts = {"dwdu": rng.normal(size=(5, 5, n)), "face_jac": rng.normal(size=(2, 5, 5, ni)),
             "bnd_jac": rng.normal(size=(5, 5, nb))}
    shift = rng.uniform(1, 2, n)
    a = jacobian.assemble(m, parts, shift)
    x = rng.normal(size=(5, n))
    dw = np.einsum("ijc,jc->ic", parts["dwdu"], x)  # W perturbation per cell
    y = shift * x
    for f in range(m.n_faces):
        o = m.owner[f]
        if f < ni:
            nb_ = m.neighbour[f]
            df = parts["face_jac"][0, :, :, f] @ dw[:, o] + parts["face_jac"][1, :, :, f] @ dw[:, nb_]
            y[:, o] += df
            y[:, nb_] -= df
        else:
            y[:, o] += parts["bnd_jac"][:, :, f - ni] @ dw[:, o]
    got = jacobian.to_components(a @ jacobian.to_cells(x, n), n)
    np.testing.assert_allclose(got, y, rtol=1e-12, atol=1e-12)
    assert a.blocksize == (5, 5) and a.shape == (5 * n, 5 * n)


class LinearStepper:
    """A stepper whose residual is linear, R(U) = A U - b with A the assembled matrix of fixed
    parts: the implicit solve must reach A U = b."""

    def __init__(self, mesh, parts, b):
        self.mesh, self.parts, self.b = mesh, parts, b
        self.a = jacobian.assemble(mesh, parts)

    def residual_and_jacobian(self, u):
        n = self.mesh.n_cells
        res = jacobian.to_components(self.a @ jacobian.to_cells(u, n), n) - self.b
        return dict(self.parts, res=res)

    def spectral_radius(self):
        return np.ones(self.mesh.n_cells)


def test_the_implicit_solve_reaches_the_solution_of_a_linear_residual():
    from scipy.sparse.linalg import spsolve

    from autan_amd import implicit

    m = meshes.kuhn_tets(2, 0.1, 0)
    n, ni, nb = m.n_cells, m.n_interior, m.n_boundary
    rng = np.random.default_rng(7)
    eye = np.eye(5)[:, :, None]
    # Upwind-like blocks (J_L ~ +I/2, J_R ~ -I/2, the boundary's +I): an M-matrix, whose
    # pseudo-time problem V dU/dt = -R is stable.
    face = np.stack([0.5 * np.repeat(eye, ni, axis=2), -0.5 * np.repeat(eye, ni, axis=2)])
    parts = {"dwdu": np.repeat(eye, n, axis=2), "face_jac": face + 0.02 * rng.normal(size=(2, 5, 5, ni)),
             "bnd_jac": np.repeat(eye, nb, axis=2) + 0.02 * rng.normal(size=(5, 5, nb))}
    b = rng.normal(size=(5, n))
    st = LinearStepper(m, parts, b)
    out = implicit.solve(st, np.zeros(5 * n), max_iterations=30, residual_orders=10.0, cfl0=1.0)
    assert out.converged and out.iterations < 30
    exact = jacobian.to_components(spsolve(st.a.tocsc(), jacobian.to_cells(b, n)), n)
    np.testing.assert_allclose(out.u.reshape(5, n), exact, rtol=1e-8, atol=1e-10)
    assert out.cfl == sorted(out.cfl)  # the CFL grows as the residual falls


def test_colours_share_no_face():
    m = meshes.kuhn_tets(3, 0.1, 0)
    cols = jacobian.colours(m)
    colour = np.full(m.n_cells, -1)
    for k, cells in enumerate(cols):
        assert np.all(colour[cells] == -1)
        colour[cells] = k
    assert np.all(colour >= 0)
    assert not np.any(colour[m.owner[: m.n_interior]] == colour[m.neighbour])


def test_gauss_seidel_reaches_the_linear_solution():
    """Symmetric multicolour block Gauss-Seidel on an M-matrix of the face structure converges to
    the exact solution."""
    from scipy.sparse.linalg import spsolve

    m = meshes.kuhn_tets(2, 0.1, 0)
    n, ni, nb = m.n_cells, m.n_interior, m.n_boundary
    rng = np.random.default_rng(9)
    eye = np.eye(5)[:, :, None]
    face = np.stack([0.5 * np.repeat(eye, ni, axis=2), -0.5 * np.repeat(eye, ni, axis=2)])
    parts = {"dwdu": np.repeat(eye, n, axis=2), "face_jac": face + 0.02 * rng.normal(size=(2, 5, 5, ni)),
             "bnd_jac": np.repeat(eye, nb, axis=2) + 0.02 * rng.normal(size=(5, 5, nb))}
    a = jacobian.assemble(m, parts, shift=np.full(n, 0.1))
    b = rng.normal(size=5 * n)
    exact = spsolve(a.tocsc(), b)
    x = jacobian.gauss_seidel(a, b, np.zeros(5 * n), 60, jacobian.colours(m))
    np.testing.assert_allclose(x, exact, rtol=0, atol=1e-9 * np.abs(exact).max())
