"""Algorithmic HBM bytes of each kernel of the step: the reference's "useful bytes" convention
(physics/bench/euler_bench.cu:321-377), every array a kernel touches counted once per launch at its
algorithmic size (padding excluded), extended to the fused kernels of this library.

Counts: N cells, Fi interior faces, Fb boundary faces, S gradient-stencil entries, I incidences
(cell-face pairs), Nbo distinct owners of boundary faces, Fm faces whose moments are not all +0
(the fused face kernels read no moment of the others). Sizes: s = sizeof(state) = sizeof(grad) =
sizeof(geom value) = 4 in fp32+acc64 and 8 in fp64; r = 8 (residual, conserved, FP64 volume);
i = 4 (indices).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Counts:
    N: int
    Fi: int
    Fb: int
    S: int
    I: int
    Nbo: int
    source: bool = True
    limiter: bool = True
    second_order: bool = True
    planar: bool = True
    Fm: int = -1  # faces with a moment not +0 (-1: every face)

    @property
    def F(self) -> int:
        return self.Fi + self.Fb

    @classmethod
    def of(cls, mesh, source=True, limiter=True, second_order=True, planar=None) -> "Counts":
        nb = mesh.n_boundary
        nbo = int(np.unique(mesh.owner[mesh.n_interior :]).size) if nb else 0
        mom = np.asarray(mesh.face_moment, dtype=np.float64).reshape(9, mesh.n_faces)
        fm = int(np.count_nonzero(np.any(mom.view(np.uint64) != 0, axis=0)))
        return cls(
            N=mesh.n_cells,
            Fi=mesh.n_interior,
            Fb=nb,
            S=mesh.n_stencil,
            I=int(mesh.cell_face_offsets[-1]),
            Nbo=nbo,
            source=source,
            limiter=limiter,
            second_order=second_order,
            planar=mesh.all_planar if planar is None else planar,
            Fm=fm,
        )


def kernel_bytes(c: Counts, policy: str = "fp32+acc64", local: bool = False) -> dict[str, float]:
    """Algorithmic bytes per launch of each kernel of the step."""
    s = 4 if policy == "fp32+acc64" else 8
    r, i = 8, 4
    N, Fi, Fb, F, S, I = c.N, c.Fi, c.Fb, c.F, c.S, c.I
    moment = 0 if c.planar else 9 * s
    out = {}
    # u -> stored state -> primitives
    out["prim"] = N * (5 * r + 5 * s)
    # boundary states: owner, code, data, normal, owner's W, W_b
    out["bstate"] = Fb * (i + 4 + 5 * s + 3 * s + 5 * s) + c.Nbo * 5 * s
    # gradient + bounds over the stencil, then the limiter over the cell's faces: W, stencil
    # lengths and entries, weights, boundary W, the limiter's volume, incidences and face offsets
    # r = x_f - x_c of each side; writes the gradient and phi.
    grad = N * (5 * s + i + 15 * s + 5 * s) + S * (i + 3 * s) + Fb * 5 * s
    if c.limiter:
        grad += N * (i + s) + I * i + (F + Fi) * 3 * s
    out["grad_limiter" if c.limiter else "grad"] = grad
    # face fluxes: cells' W, gradient, phi; face ends, normal + area, r of each side, moments,
    # boundary codes and data; writes one flux per face.
    recon = (15 * s + 5 * s) if c.second_order else 0
    out["flux"] = (
        N * (5 * s + recon)
        + Fi * 2 * i
        + Fb * i
        + F * (4 * s + 3 * s + moment)
        + Fi * 3 * s
        + Fb * (4 + 5 * s)
        + F * 5 * s
    )
    # gather: face offsets, incidences, fluxes, volume, source; writes res.
    out["gather"] = N * (i + s + (5 * s if c.source else 0) + 5 * r) + I * i + F * 5 * s
    out["norms"] = N * 5 * r
    # spectral radius + dt: W, face offsets, incidences, normal + area, face ends, W_b; writes
    # lambda (and the local dt); the FP64 volume.
    out["spectral"] = N * (5 * s + i + s + r + (r if local else 0)) + I * i + F * 4 * s + Fi * 2 * i + Fb * 5 * s
    # RK stages: u, stage (not on the first), R, FP64 volume, local dt; writes the stage and,
    # before the next residual, its primitives.
    base = N * (5 * r + 5 * r + r + (r if local else 0) + 5 * r)
    out["rk_prim"] = base + N * (5 * r / 2 + 5 * s)  # stage 1 reads no separate stage; averaged
    out["rk_check"] = base + N * 5 * r
    out["clear"] = 5 * 71 * 8
    # The fused face kernel: the flux kernel's inputs and the gather's, the fluxes staying on chip
    # (LDS); the first residual of a step also writes lambda (and the local dt) from the FP64 volume
    # and adds the norms of the residual it writes.
    fm = F if c.Fm < 0 else c.Fm
    out["faces"] = out["flux"] - F * 5 * s + out["gather"] - F * 5 * s - (F - fm) * moment
    out["faces_first"] = out["faces"] + N * (r + s + (r if local else 0))
    # The face kernel that also updates the stage (no residual written): u and the stage read, the
    # new stage written, the FP64 volume and local dt read; then the new stage's primitives (on the
    # last stage, the new state's: the next step's first). On the first stage the stage read is u
    # itself and the dt is the kernel's own.
    rk_io = N * (5 * r + 5 * r + 5 * r + r + (r if local else 0))
    out["faces_rk"] = out["faces"] - N * 5 * r + rk_io + N * 5 * s
    out["faces_rk_check"] = out["faces"] - N * 5 * r + rk_io + N * 5 * s  # and the next step's primitives
    out["faces_first_rk"] = out["faces_first"] - N * 5 * r + N * (5 * r + 5 * r) + N * 5 * s
    return out


def step_bytes(c: Counts, policy: str = "fp32+acc64", local: bool = False) -> float:
    """Algorithmic bytes of one SSP-RK3 step with this library's kernels (three residuals)."""
    k = kernel_bytes(c, policy, local)
    residual = k["bstate"] + k.get("grad_limiter", k.get("grad", 0)) * (1 if c.second_order else 0) + k["flux"] + k["gather"]
    return k["prim"] + 3 * residual + k["norms"] + k["spectral"] + 2 * k["rk_prim"] + k["rk_check"]


def reference_step_bytes(c: Counts, policy: str = "fp32+acc64", local: bool = False) -> float:
    """B_step of the reference's own loop structure (useful bytes per unfused loop): the
    denominator under which the reference would state its roofline fraction for this step."""
    s = g = v = 4 if policy == "fp32+acc64" else 8
    r, i, o, k, cc = 8, 4, 8, 8, 8
    N, Fi, Fb, F, S, I, Nbo = c.N, c.Fi, c.Fb, c.F, c.S, c.I, c.Nbo
    nonplanar = 0 if c.planar else 1
    b_prim = 10 * s * N
    b_bst = Fb * (i + k + 5 * s + 3 * v + 5 * s) + 5 * s * Nbo
    b_grad = 5 * s * N + 5 * s * Fb + i * S + o * N + 3 * v * S + (15 * g + 10 * s) * N
    b_lim = (5 * s + 15 * g + 10 * s + v + 3 * cc + 5 * s + o) * N + i * I + 3 * cc * F
    b_src = (5 * s + v + 5 * r) * N
    b_face = (25 * g + 3 * cc + 10 * r) * N + Fi * (2 * i + 3 * v + v + 3 * cc + 9 * v * nonplanar)
    b_bnd = Fb * (i + k + 5 * s + 3 * v + v + 3 * cc + 9 * v * nonplanar) + Nbo * (25 * g + 3 * cc + 10 * r)
    b_res = b_prim + b_bst + b_grad + (b_lim if c.limiter else 0) + b_src + b_face + b_bnd
    b_lam = 5 * s * (N + Fb) + i * I + o * N + 4 * v * F + s * N
    return (
        3 * (b_res + (40 + 5 * s) * N + (40 + 5 * r + 8 + 8 * (1 if local else 0) + 80) * N)
        + 5 * r * N
        + b_lam
        + (8 + s + 8) * N
        + (40 + 1) * N
    )
