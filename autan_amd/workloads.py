"""The synthetic workloads of the benchmark: each builds a mesh, a scheme, a state and boundary
data of the reference's benchmark configurations, with fixed seeds.

- vortex: Shu's isentropic vortex (strength 5, freestream Mach 0.5 along x) on n^3 periodic
  hexahedra of [-10, 10]^3 (the unit cube mapped by x' = 20 x - 10, so the vortex's tail at the
  period is below 1e-20), first-order Roe, fp64, global time step at CFL 0.5.
- tets: the unit cube in n^3 cubes of 6 Kuhn tetrahedra, interior nodes jittered by +-10 % of the
  spacing (seed 0), the manufactured solution with its volume source and exact `state` boundary
  faces; second-order HLLC, face2 least squares, Venkatakrishnan K = 5, global CFL 0.5.
- bump: the 3 x 1 x 1 channel with a Gaussian bump on the lower wall (height h, width w), hybrid:
  a prism layer on the wall, a band of merged 2 x 2 x 2 polyhedra with non-planar 8-node faces,
  hexahedra elsewhere; characteristic farfield inflow and outflow, slip walls and side planes,
  freestream start; second order, local time steps at CFL 0.8 (the steady device path).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from autan_amd import mesh as meshes
from autan_amd import mms
from autan_amd.stepper import Scheme, freestream_scheme

GAMMA = 1.4


@dataclass
class Workload:
    name: str
    mesh: meshes.Mesh
    scheme: Scheme
    policy: str
    u: np.ndarray  # 5 x n conserved, FP64
    bc_kind: np.ndarray  # float codes, one per boundary face
    bc: np.ndarray  # 5 x nb
    source: np.ndarray | None
    cfl: float
    local: bool
    description: str
    counts: dict = field(default_factory=dict)  # traffic.Counts.of keyword arguments


def conserved(rho, u, v, w, p, gamma=GAMMA) -> np.ndarray:
    return np.stack([rho, rho * u, rho * v, rho * w, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def vortex_state(x: np.ndarray, mach=0.5, beta=5.0, gamma=GAMMA) -> np.ndarray:
    """Shu's isentropic vortex at the origin with freestream speed of sound 1 (the reference's
    oracle, vv/src/autan_vv/oracles/inviscid.py:105-181), independent of z."""
    dx, dy = x[:, 0], x[:, 1]
    r2 = dx * dx + dy * dy
    t = 1.0 - (gamma - 1.0) * beta**2 / (8.0 * math.pi**2) * np.exp(1.0 - r2)
    bump = beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2))
    rho = t ** (1.0 / (gamma - 1.0))
    return conserved(rho, mach - bump * dy, bump * dx, np.zeros_like(rho), rho * t / gamma, gamma)


def vortex(n: int = 32, policy: str = "fp64", flux: str = "roe", order: str = "rcm") -> Workload:
    m = meshes.periodic_hex(n, -10.0, 10.0, order=order)
    sc = freestream_scheme(1.0, (0.5, 0.0, 0.0), 1.0 / GAMMA, flux=flux, second_order=False, limiter=False)
    sc.gas_constant = 1.0 / GAMMA
    u = vortex_state(m.centroids()).reshape(-1)
    return Workload(
        "vortex", m, sc, policy, u, np.zeros(0), np.zeros(0), None, 0.5, False,
        f"isentropic vortex (strength 5, Mach 0.5) on {n}^3 periodic hexahedra of [-10, 10]^3, first-order "
        f"{flux.upper()}, {policy}, SSP-RK3 CFL 0.5, {order} cell order",
        dict(source=False, limiter=False, second_order=False),
    )


def tets(n: int = 128, policy: str = "fp32+acc64", flux: str = "hllc", order: str = "morton") -> Workload:
    m = meshes.kuhn_tets(n, 0.1, 0, order=order)
    kinds = np.zeros(m.n_boundary, dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    rho, vel, p = mms.freestream()
    sc = freestream_scheme(rho, vel, p, flux=flux, second_order=True, limiter=True)
    sc.venkat_k = 5.0
    x = m.centroids()
    u = mms.conserved(mms.primitive(x)).reshape(-1)
    src = mms.source(x).reshape(-1)
    bc = mms.primitive(m.face_centroids()[m.n_interior :]).reshape(-1)
    return Workload(
        "tets", m, sc, policy, u, kinds.astype(np.float64), bc, src, 0.5, False,
        f"Kuhn tetrahedra {n}^3 x 6 (jitter 10 %, seed 0), manufactured solution, second-order {flux.upper()}, "
        f"face2 LSQ + Venkatakrishnan K=5, {policy}, SSP-RK3 CFL 0.5, {order} cell order",
        dict(source=True, limiter=True, second_order=True),
    )


def bump_freestream(mach: float = 0.5, alpha_deg: float = 0.0) -> np.ndarray:
    """Primitive freestream (rho, u, v, w, p) with rho = 1 and speed of sound 1."""
    a = math.radians(alpha_deg)
    return np.array([1.0, mach * math.cos(a), mach * math.sin(a), 0.0, 1.0 / GAMMA])


def bump_mesh(nx: int = 462, ny: int = 154, nz: int = 154, h: float = 0.1, w: float = 0.2, prism_layers: int = 8,
              order: str = "morton") -> meshes.Mesh:
    return meshes.bump_hybrid(nx, ny, nz, h=h, w=w, prism_layers=prism_layers, order=order)


def bump(nx: int = 462, ny: int = 154, nz: int = 154, policy: str = "fp32+acc64", flux: str = "hllc",
         h: float = 0.1, w: float = 0.2, mach: float = 0.5, alpha_deg: float = 0.0, prism_layers: int = 8,
         order: str = "morton", mesh: meshes.Mesh | None = None) -> Workload:
    m = mesh if mesh is not None else bump_mesh(nx, ny, nz, h, w, prism_layers, order)
    wf = bump_freestream(mach, alpha_deg)
    # zones 0..5: xmin, xmax (farfield), ymin (the bump wall), ymax, zmin, zmax (slip)
    by_zone = [meshes.FARFIELD, meshes.FARFIELD, meshes.SLIP, meshes.SLIP, meshes.SLIP, meshes.SLIP]
    kinds = np.array(by_zone, dtype=np.int32)[m.face_zone]
    meshes.set_boundary_kinds(m, kinds)
    sc = freestream_scheme(wf[0], wf[1:4], wf[4], flux=flux, second_order=True, limiter=True)
    bc = np.repeat(wf[:, None], m.n_boundary, axis=1).reshape(-1)
    n = m.n_cells
    u = conserved(np.full(n, wf[0]), np.full(n, wf[1]), np.full(n, wf[2]), np.full(n, wf[3]), np.full(n, wf[4]))
    return Workload(
        "bump", m, sc, policy, u.reshape(-1), kinds.astype(np.float64), bc, None, 0.8, True,
        f"hybrid bump channel 3x1x1 (h={h}, w={w}; {nx}x{ny}x{nz} lattice, {prism_layers} prism rows, merged "
        f"polyhedral band): {n} cells, Mach {mach}, alpha {alpha_deg} deg, farfield + slip walls, second-order "
        f"{flux.upper()} + Venkatakrishnan K={sc.venkat_k:g}, {policy}, SSP-RK3 local time steps CFL 0.8, "
        f"{order} cell order",
        dict(source=False, limiter=True, second_order=True),
    )


def wall_faces(m: meshes.Mesh) -> np.ndarray:
    """The bump channel's lower wall (boundary zone 2, ymin), as global face indices."""
    return (np.nonzero(m.face_zone == 2)[0] + m.n_interior).astype(np.int32)
