"""An implicit steady solve on the stepper's residual and approximate Jacobian: pseudo-transient
continuation, (V / dt_c I + dR/dU) dU = -R with the pseudo-time step of each cell dt_c = CFL V_c /
lambda_c (lambda_c the convective spectral radius, so V / dt_c = lambda_c / CFL), the CFL grown as
the residual falls (switched evolution relaxation: CFL_k = CFL_0 |R_0| / |R_k|, capped), and each
linear system solved on the host: by sparse LU, by GMRES with block-Jacobi preconditioning, or by a
few symmetric multicolour block Gauss-Seidel sweeps, the reference's relaxation.

The Jacobian is the reference's (Euler::residual_and_jacobian and Euler::assemble,
physics/src/euler.cpp:983-1053); its implicit solvers iterate on the device (block Gauss-Seidel,
line smoothers, multigrid), which this host solve stands in for on meshes small enough to factor."""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from autan_amd import jacobian


@dataclass
class ImplicitResult:
    converged: bool
    iterations: int  # linear solves done
    residuals: np.ndarray  # (iterations + 1, 5): RMS of R per equation before each solve and at the end
    cfl: list[float]
    u: np.ndarray  # the final state, 5 x n_cells flat


def rms(res: np.ndarray) -> np.ndarray:
    r = np.asarray(res, dtype=np.float64).reshape(5, -1)
    return np.sqrt((r * r).sum(axis=1) / r.shape[1])


def block_jacobi(a, n: int):
    """The inverse of the matrix's 5 x 5 diagonal blocks, as a preconditioner (scipy LinearOperator)."""
    from scipy.sparse.linalg import LinearOperator

    inv = np.linalg.inv(jacobian.diagonal_blocks(a, n))
    return LinearOperator((5 * n, 5 * n), matvec=lambda x: np.einsum("cij,cj->ci", inv, x.reshape(n, 5)).reshape(-1))


def solve(stepper, u0: np.ndarray, max_iterations: int = 50, residual_orders: float = 8.0, cfl0: float = 10.0,
          cfl_max: float = 1e10, linear: str = "lu", linear_rtol: float = 1e-3, sweeps: int = 4) -> ImplicitResult:
    """Steps from u0 until every equation's RMS residual has dropped `residual_orders` orders from
    its first value (equations whose first value is zero are met), or `max_iterations` solves.
    `stepper` gives residual_and_jacobian(u) and spectral_radius() (after a residual) and has
    `mesh`; a DeviceStepper does. `linear`: "lu" factors each system exactly; "gmres" solves it to
    `linear_rtol` by GMRES preconditioned with the inverse diagonal blocks (block Jacobi); "gs" takes
    `sweeps` symmetric multicolour block Gauss-Seidel sweeps from zero, the reference's relaxation
    (physics/include/autan/physics/block_matrix.hpp, BlockMatrix::gauss_seidel)."""
    from scipy.sparse.linalg import gmres, spsolve

    if linear not in ("lu", "gmres", "gs"):
        raise ValueError("linear is 'lu', 'gmres' or 'gs'")

    mesh = stepper.mesh
    n = mesh.n_cells
    cols = jacobian.colours(mesh) if linear == "gs" else None
    u = np.asarray(u0, dtype=np.float64).reshape(-1).copy()
    history, cfls = [], []
    first = None
    for it in range(max_iterations + 1):
        parts = stepper.residual_and_jacobian(u)
        r = rms(parts["res"])
        history.append(r)
        if first is None:
            first = r
        met = np.all((r <= first * 10.0 ** (-residual_orders)) | (first == 0.0))
        if met or it == max_iterations or not np.all(np.isfinite(r)):
            return ImplicitResult(bool(met), it, np.array(history), cfls, u)
        total0, total = math.sqrt(float(first @ first)), math.sqrt(float(r @ r))
        cfl = min(cfl_max, cfl0 * total0 / max(total, 1e-300))
        cfls.append(cfl)
        lam = np.asarray(stepper.spectral_radius(), dtype=np.float64).reshape(-1)
        a = jacobian.assemble(mesh, parts, shift=lam / cfl)
        rhs = -jacobian.to_cells(parts["res"], n)
        if linear == "lu":
            du = spsolve(a.tocsc(), rhs)
        elif linear == "gs":
            du = jacobian.gauss_seidel(a, rhs, np.zeros_like(rhs), sweeps, cols)
        else:
            du, _ = gmres(a.tocsr(), rhs, rtol=linear_rtol, restart=50, maxiter=20, M=block_jacobi(a, n))
        u = u + jacobian.to_components(du, n).reshape(-1)
    raise AssertionError("unreachable")
