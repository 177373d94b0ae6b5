"""The residual's Jacobian as a block sparse matrix from the parts DeviceStepper.residual_and_jacobian
gives (Euler::assemble, physics/src/euler.cpp:993-1053): per cell a 5 x 5 diagonal block
shift_c I + sum over its faces of +-J_side D_c, and per interior face the two off-diagonal blocks
A(owner, neighbour) = +J_R D_neighbour and A(neighbour, owner) = -J_L D_owner, with R_owner += F and
R_neighbour -= F. Rows and columns are cell-major (5 c + k); `to_cells` and `to_components` convert
from and to the stepper's component-major arrays (k n + c)."""

from __future__ import annotations

import numpy as np


def to_cells(x: np.ndarray, n: int) -> np.ndarray:
    """Component-major 5 x n (flat or not) to cell-major (5 n,)."""
    return np.asarray(x, dtype=np.float64).reshape(5, n).T.reshape(-1)


def to_components(x: np.ndarray, n: int) -> np.ndarray:
    """Cell-major (5 n,) to component-major (5, n)."""
    return np.asarray(x, dtype=np.float64).reshape(n, 5).T


def assemble(mesh, parts: dict, shift: np.ndarray | None = None):
    """The block matrix (scipy.sparse BSR, 5 x 5 blocks) of shift I + dR/dU for the mesh and the
    Jacobian parts (dwdu, face_jac, bnd_jac) of one state."""
    import scipy.sparse as sp

    n, ni = mesh.n_cells, mesh.n_interior
    dm = np.moveaxis(parts["dwdu"], 2, 0)  # (n, 5, 5)
    jl = np.moveaxis(parts["face_jac"][0], 2, 0)  # (ni, 5, 5)
    jr = np.moveaxis(parts["face_jac"][1], 2, 0)
    jb = np.moveaxis(parts["bnd_jac"], 2, 0)  # (nb, 5, 5)
    own = np.asarray(mesh.owner[:ni], dtype=np.int64)
    nei = np.asarray(mesh.neighbour, dtype=np.int64)
    bown = np.asarray(mesh.owner[ni:], dtype=np.int64)
    left = jl @ dm[own]  # J_L D_owner
    right = jr @ dm[nei]  # J_R D_neighbour
    diag = np.zeros((n, 5, 5))
    if shift is not None:
        diag += np.asarray(shift, dtype=np.float64)[:, None, None] * np.eye(5)
    np.add.at(diag, own, left)
    np.add.at(diag, nei, -right)
    np.add.at(diag, bown, jb @ dm[bown])
    rows = np.concatenate([np.arange(n), own, nei])
    cols = np.concatenate([np.arange(n), nei, own])
    blocks = np.concatenate([diag, right, -left])
    i = np.arange(5)
    r = (5 * rows[:, None, None] + i[None, :, None]) + 0 * i[None, None, :]
    c = (5 * cols[:, None, None] + i[None, None, :]) + 0 * i[None, :, None]
    a = sp.coo_matrix((blocks.reshape(-1), (r.reshape(-1), c.reshape(-1))), shape=(5 * n, 5 * n))
    return a.tobsr(blocksize=(5, 5))


def colours(mesh) -> list[np.ndarray]:
    """Cells in colours that share no face (greedy, in cell order: each cell the least colour none
    of its face neighbours has taken)."""
    n, ni = mesh.n_cells, mesh.n_interior
    own = np.asarray(mesh.owner[:ni], dtype=np.int64)
    nei = np.asarray(mesh.neighbour, dtype=np.int64)
    adj = [[] for _ in range(n)]
    for a, b in zip(own.tolist(), nei.tolist()):
        adj[a].append(b)
        adj[b].append(a)
    colour = np.full(n, -1, dtype=np.int64)
    for c in range(n):
        taken = {int(colour[d]) for d in adj[c] if colour[d] >= 0}
        k = 0
        while k in taken:
            k += 1
        colour[c] = k
    return [np.flatnonzero(colour == k) for k in range(int(colour.max()) + 1)]


def diagonal_blocks(a, n: int) -> np.ndarray:
    """The matrix's 5 x 5 diagonal blocks, (n, 5, 5)."""
    bsr = a.tobsr(blocksize=(5, 5))
    d = np.zeros((n, 5, 5))
    rows = np.repeat(np.arange(n), np.diff(bsr.indptr))
    on_diag = bsr.indices == rows
    d[rows[on_diag]] = bsr.data[on_diag]
    return d


def gauss_seidel(a, b: np.ndarray, x: np.ndarray, sweeps: int, cols: list[np.ndarray]) -> np.ndarray:
    """`sweeps` symmetric multicolour block Gauss-Seidel sweeps on A x = b from x (cell-major
    vectors), the colours in increasing then decreasing order, after the diagonal blocks are
    inverted (BlockMatrix::factor and gauss_seidel, physics/include/autan/physics/block_matrix.hpp):
    the cells of one colour share no face, so they update together, each from its row with the
    other cells' latest values."""
    n = a.shape[0] // 5
    x = np.asarray(x, dtype=np.float64).copy()
    csr = a.tocsr()
    d = diagonal_blocks(a, n)
    inv = np.linalg.inv(d)
    order = list(cols) + list(reversed(cols))
    for _ in range(sweeps):
        for cells in order:
            ax = (csr @ x).reshape(n, 5)
            xc = x.reshape(n, 5)
            r = b.reshape(n, 5)[cells] - ax[cells] + np.einsum("cij,cj->ci", d[cells], xc[cells])
            xc[cells] = np.einsum("cij,cj->ci", inv[cells], r)
    return x
