"""Test cases shared by the CPU and GPU tests: meshes, schemes, states and boundary data.

The flows are smooth analytic fields about a freestream, so every residual is well defined; the
boundary data follow the BoundaryKind contract of physics/include/autan/physics/euler.hpp:30-37.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from autan_amd import mesh as meshes
from autan_amd.stepper import Scheme, freestream_scheme

GOLDEN = Path(__file__).resolve().parent / "golden"
GAMMA = 1.4
RGAS = 287.0


@dataclass
class Case:
    name: str
    mesh: meshes.Mesh
    scheme: Scheme
    u: np.ndarray  # 5 x n conserved, FP64
    bc_kind: np.ndarray
    bc: np.ndarray  # 5 x nb
    source: np.ndarray | None = None
    uses_exp_log: bool = False  # farfield or inlet faces: device exp/log differ from glibc's


def conserved(rho, u, v, w, p, gamma=GAMMA) -> np.ndarray:
    return np.stack([rho, rho * u, rho * v, rho * w, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def vortex_state(x: np.ndarray, mach=0.5, beta=5.0, gamma=GAMMA) -> np.ndarray:
    """Shu's isentropic vortex at the origin, freestream speed of sound 1 (restated from the
    reference's oracle, vv/src/autan_vv/oracles/inviscid.py:105-181), z-independent."""
    dx, dy = x[:, 0], x[:, 1]
    r2 = dx * dx + dy * dy
    t = 1.0 - (gamma - 1.0) * beta**2 / (8.0 * math.pi**2) * np.exp(1.0 - r2)
    bump = beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2))
    rho = t ** (1.0 / (gamma - 1.0))
    return conserved(rho, mach - bump * dy, bump * dx, np.zeros_like(rho), rho * t / gamma, gamma)


def smooth_state(x: np.ndarray, rho0=1.0, vel=(0.3, 0.1, 0.05), p0=1.0 / GAMMA, amp=0.05, k=2.0) -> np.ndarray:
    """A smooth perturbation of a uniform flow (sinusoids of amplitude amp)."""
    s = np.sin(k * x[:, 0] + 0.3) * np.cos(k * x[:, 1] - 0.2) * np.sin(k * x[:, 2] + 0.7)
    c = np.cos(k * x[:, 0] - 0.5) * np.sin(k * x[:, 1] + 0.1)
    rho = rho0 * (1.0 + amp * s)
    u = vel[0] + amp * c
    v = vel[1] - amp * s
    w = vel[2] + amp * c * s
    p = p0 * (1.0 + amp * c)
    return conserved(rho, u, v, w, p)


def boundary_data(kinds: np.ndarray, w_free: np.ndarray, gamma=GAMMA, rgas=RGAS) -> np.ndarray:
    """Data of each boundary face for its condition (5 x nb): the freestream W for state and
    farfield faces, (p0, T0) for inlets, p for outlets, an isothermal wall at 1.1 T."""
    nb = kinds.size
    rho, u, v, w, p = w_free
    t = p / (rho * rgas)
    m2 = (u * u + v * v + w * w) * rho / (gamma * p)
    t0 = t * (1.0 + 0.5 * (gamma - 1.0) * m2)
    p0 = p * (t0 / t) ** (gamma / (gamma - 1.0))
    out = np.zeros((5, nb))
    for b, k in enumerate(kinds):
        if k in (meshes.STATE, meshes.FARFIELD):
            out[:, b] = w_free
        elif k == meshes.INLET:
            out[:2, b] = (p0, t0)
        elif k == meshes.OUTLET:
            out[4, b] = 0.98 * p
        elif k == meshes.WALL:
            out[0, b] = 1.1 * t if b % 2 == 0 else 0.0
    return out


def state_faces_data(m: meshes.Mesh, field) -> np.ndarray:
    """BC `state` data from a primitive field evaluated at the boundary face centroids."""
    xf = m.face_centroids()[m.n_interior :]
    return field(xf)


def _prim(u: np.ndarray, gamma=GAMMA) -> np.ndarray:
    rho = u[0]
    vel = u[1:4] / rho
    p = (gamma - 1.0) * (u[4] - 0.5 * rho * (vel * vel).sum(axis=0))
    return np.stack([rho, vel[0], vel[1], vel[2], p])


def vortex_case(n=16, flux="roe", second_order=False, limiter=False, order="rcm") -> Case:
    """The periodic isentropic vortex: n^3 hexahedra of [-10, 10]^3 (the unit cube mapped by
    x' = 20 x - 10), freestream Mach 0.5 along x, strength 5."""
    m = meshes.periodic_hex(n, -10.0, 10.0, order=order)
    sc = freestream_scheme(1.0, (0.5, 0.0, 0.0), 1.0 / GAMMA, flux=flux, second_order=second_order, limiter=limiter)
    sc.gas_constant = 1.0 / GAMMA
    u = vortex_state(m.centroids())
    return Case("vortex", m, sc, u.reshape(-1), np.zeros(0), np.zeros((5, 0)))


def tets_case(n=6, flux="hllc", second_order=True, limiter=True, venkat_k=5.0, order="rcm", source=True) -> Case:
    """Kuhn tetrahedra of the jittered unit cube, every boundary face a `state` face holding the
    exact field, with a volume source."""
    m = meshes.kuhn_tets(n, 0.1, 0, order=order)
    w_free = np.array([1.0, 0.3 * math.sqrt(GAMMA * (1 / GAMMA) / 1.0), 0.0, 0.0, 1.0 / GAMMA])
    sc = freestream_scheme(1.0, w_free[1:4], w_free[4], flux=flux, second_order=second_order, limiter=limiter)
    sc.venkat_k = venkat_k
    x = m.centroids()
    u = smooth_state(x, vel=(w_free[1], 0.05, -0.02))
    kinds = np.zeros(m.n_boundary, dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    bc = _prim(smooth_state(m.face_centroids()[m.n_interior :], vel=(w_free[1], 0.05, -0.02)))
    src = None
    if source:
        src = 0.01 * np.stack([np.sin(3 * x[:, 0] + k) * np.cos(2 * x[:, 1] - k) for k in range(5)])
    return Case("tets", m, sc, u.reshape(-1), kinds.astype(np.float64), bc, src)


def fixture_case(name: str, flux="ausm+up", second_order=True, limiter=True, kinds_by_zone=None, order=None) -> Case:
    """A reference mesh fixture (OpenFOAM polyMesh under tests/golden/meshes) with a condition per
    boundary zone."""
    m = meshes.read_polymesh(GOLDEN / "meshes" / name, order=order)
    w_free = np.array([1.0, 0.4, 0.1, -0.05, 1.0 / GAMMA])
    sc = freestream_scheme(1.0, w_free[1:4], w_free[4], flux=flux, second_order=second_order, limiter=limiter)
    if kinds_by_zone is None:
        kinds_by_zone = [meshes.FARFIELD, meshes.OUTLET, meshes.SLIP, meshes.WALL, meshes.INLET, meshes.STATE]
    zone = m.face_zone
    kinds = np.array([kinds_by_zone[z % len(kinds_by_zone)] for z in zone], dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    bc = boundary_data(kinds, w_free)
    u = smooth_state(m.centroids(), vel=tuple(w_free[1:4]))
    uses = bool(np.isin(kinds, [meshes.FARFIELD, meshes.INLET]).any())
    return Case(name, m, sc, u.reshape(-1), kinds.astype(np.float64), bc, None, uses)


def bump_case(nx=24, ny=8, nz=4, flux="ausm+up", second_order=True, limiter=True, hybrid=False) -> Case:
    """The bump channel, in hexahedra or hybrid (prism wall layer, band of merged polyhedra):
    farfield inflow and outflow, slip walls top and bottom, slip side planes."""
    m = meshes.bump_hybrid(nx, ny, nz) if hybrid else meshes.bump_hex(nx, ny, nz)
    mach = 0.5
    c = 1.0
    w_free = np.array([1.0, mach * c, 0.0, 0.0, 1.0 / GAMMA])
    sc = freestream_scheme(1.0, w_free[1:4], w_free[4], flux=flux, second_order=second_order, limiter=limiter)
    kinds_by_zone = [meshes.FARFIELD, meshes.FARFIELD, meshes.SLIP, meshes.SLIP, meshes.SLIP, meshes.SLIP]
    kinds = np.array([kinds_by_zone[z] for z in m.face_zone], dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    bc = boundary_data(kinds, w_free)
    u = smooth_state(m.centroids(), vel=tuple(w_free[1:4]), amp=0.02)
    return Case("bump-hybrid" if hybrid else "bump", m, sc, u.reshape(-1), kinds.astype(np.float64), bc, None, True)


def normwise_units(case: Case, policy_eps: float) -> np.ndarray:
    """The reference's per-cell tolerance unit eps f_k A_i (tests/physics/test_euler_device.cpp:
    68-211): f_k = (rho q, rho q^2 x 3, rho H q) of the freestream with q = |v| + c, A_i the cell's
    surface. Shape (5, n)."""
    m = case.mesh
    ref = np.array(case.scheme.reference, dtype=float)
    g = case.scheme.gamma
    rho = ref[0]
    v = ref[1:4] / rho
    p = (g - 1.0) * (ref[4] - 0.5 * rho * (v @ v))
    c = math.sqrt(g * p / rho)
    q = math.sqrt(v @ v) + c
    h = c * c / (g - 1.0) + 0.5 * (v @ v)
    f = np.array([rho * q, rho * q * q, rho * q * q, rho * q * q, rho * h * q])
    area = np.zeros(m.n_cells)
    np.add.at(area, m.owner, m.face_area)
    np.add.at(area, m.neighbour, m.face_area[: m.n_interior])
    return policy_eps * f[:, None] * area[None, :]


class OracleTarget:
    """The oracle as the stepper of a steady run (step / evaluate / set_second_order), holding the
    state u: the CPU side of the steady-run comparisons."""

    def __init__(self, ora, u) -> None:
        self.ora = ora
        self.u = np.array(u, dtype=np.float64).reshape(-1)
        self.n = self.u.size // 5

    def norms(self, res: np.ndarray) -> np.ndarray:
        from oracle import oracle as orc

        r = res.reshape(5, self.n)
        return np.array([math.sqrt(orc.exact_sum(r[k] * r[k]) / self.n) for k in range(5)])

    def step(self, cfl):
        self.u, norms, _h, inv = self.ora.step(self.u, cfl=cfl, local=True)
        return norms, inv

    def evaluate(self):
        res = self.ora.residual(self.u)
        return self.norms(res), self.ora.first_invalid(self.u, res)

    def set_second_order(self, on: bool) -> None:
        self.ora.set_second_order(on)


class OracleImplicitTarget(OracleTarget):
    """The oracle as the stepper of a steady implicit run (evaluate / set_second_order / jacobian /
    solve): R(u) and the Jacobian's parts from oracle_residual_and_jacobian, the spectral radius,
    then oracle_implicit_solve's assembly, GCR, sweeps and admissible update of u."""

    def __init__(self, ora, u, mesh, gamma) -> None:
        super().__init__(ora, u)
        self.mesh, self.gamma = mesh, gamma
        self.parts = None
        self.second = True
        self.second_of_parts = True

    def set_second_order(self, on: bool) -> None:
        super().set_second_order(on)
        self.second = bool(on)

    def jacobian(self):
        self.parts = self.ora.residual_and_jacobian(self.u)
        self.second_of_parts = self.second
        return self.norms(self.parts["res"]), self.ora.first_invalid(self.u, self.parts["res"])

    def solve(self, cfl):
        from oracle import oracle as orc

        lam = self.ora.spectral_radius()
        self.u, self.du, rep, _ = orc.implicit_solve(self.mesh, self.parts, lam, cfl, self.gamma, self.u, oracle=self.ora,
                                                     second=self.second_of_parts)
        return [rep["cycles"], rep["residual_drop"], rep["alpha"], float(rep["failed"]), float(rep["singular"])]
