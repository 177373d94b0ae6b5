"""ctypes wrapper of the CPU oracle (oracle/euler_oracle.cpp): test infrastructure only.

`Oracle(mesh, scheme, policy, ...)` takes the same arrays as the device stepper and evaluates the
reference's residual, SSP-RK3 step, norms and validity on the host in one thread.
"""

from __future__ import annotations

import ctypes
import functools
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"

_d = ctypes.c_double
_i32 = ctypes.c_int32
_p = ctypes.c_void_p


class _Scheme(ctypes.Structure):
    _fields_ = [
        ("flux", ctypes.c_int),
        ("second_order", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("planar", ctypes.c_int),
        ("gamma", _d),
        ("gas_constant", _d),
        ("mach_ref", _d),
        ("venkat_k", _d),
        ("length", _d),
        ("scale", _d * 5),
        ("reference", _d * 5),
    ]


class _Inputs(ctypes.Structure):
    _fields_ = [
        ("n_cells", _i32),
        ("n_faces", _i32),
        ("n_interior", _i32),
        ("owner", _p),
        ("neighbour", _p),
        ("cell_face_offsets", _p),
        ("cell_faces", _p),
        ("face_normal", _p),
        ("face_area", _p),
        ("r_left", _p),
        ("r_right", _p),
        ("face_moment", _p),
        ("cell_volume", _p),
        ("stencil_offsets", _p),
        ("stencil", _p),
        ("stencil_weights", _p),
        ("bc_kind", _p),
        ("bc", _p),
        ("source", _p),
    ]


@functools.cache
def lib() -> ctypes.CDLL:
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is not built: run autan_amd/build.py oracle")
    L = ctypes.CDLL(str(LIB))
    L.oracle_exact_sum.restype = _d
    L.oracle_exact_sum.argtypes = [_p, ctypes.c_int64]
    L.oracle_venkat.restype = _d
    L.oracle_venkat.argtypes = [_d, _d, _d, _d]
    L.oracle_destroy.argtypes = [_p]
    L.oracle_residual.argtypes = [_p, _p, _p, ctypes.c_int]
    L.oracle_spectral_radius.argtypes = [_p, _p]
    L.oracle_face_fluxes.argtypes = [_p, _p]
    L.oracle_arrays.argtypes = [_p] * 7
    L.oracle_set_second_order.argtypes = [_p, ctypes.c_int]
    L.oracle_first_invalid.argtypes = [_p, _p, _p]
    L.oracle_step.argtypes = [_p, _p, _d, ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _p, _p, _p]
    L.oracle_flux.argtypes = [ctypes.c_int, ctypes.c_int, _p, _p, _p, _d, _d, _d, _p]
    L.oracle_physical_flux.argtypes = [_p, _p, _d, _d, _p]
    L.oracle_boundary_value.argtypes = [_p, _d, _p, _p, _d, _d, _d, _p]
    L.oracle_to_primitive.argtypes = [_p, _p, _d, _p]
    L.oracle_face_jacobian.argtypes = [ctypes.c_int, ctypes.c_int, _p, _p, _p, _d, _d, _d, _p]
    L.oracle_boundary_jacobian.argtypes = [ctypes.c_int, ctypes.c_int, _p, _d, _p, _p, _d, _d, _d, _d, _p]
    L.oracle_dwdu.argtypes = [ctypes.c_int, _p, _p, _d, _p]
    L.oracle_warp.argtypes = [_p] * 6 + [_d, _d, _p]
    L.oracle_set_threads.argtypes = [ctypes.c_int]
    L.oracle_face_force.argtypes = [_p, _p, _p, ctypes.c_int64, _p]
    L.oracle_conservation.argtypes = [_p, _p, _p]
    L.oracle_threads.restype = ctypes.c_int
    L.oracle_residual_and_jacobian.argtypes = [_p] * 6
    L.oracle_admissible_step.restype = _d
    L.oracle_admissible_step.argtypes = [ctypes.c_int64, _p, _p, _d]
    L.oracle_implicit_solve.argtypes = [_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64] + [_p] * 8 + [_d, _p, _d, _p, _p, _p, _p]
    return L


def set_threads(n: int) -> None:
    """Threads of the gather loops (1: the serial loops of the reference's CPU path). Results do
    not depend on the count."""
    lib().oracle_set_threads(int(n))


def threads() -> int:
    return lib().oracle_threads()


def _a(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64)


def exact_sum(x: np.ndarray) -> float:
    x = _a(x).reshape(-1)
    return lib().oracle_exact_sum(x.ctypes.data, x.size)


def flux(kind: int, wl, wr, n, gamma=1.4, minf=0.5, pref=0.0, policy=0) -> np.ndarray:
    out = np.empty(5)
    wl, wr, n = _a(wl), _a(wr), _a(n)
    lib().oracle_flux(policy, kind, wl.ctypes.data, wr.ctypes.data, n.ctypes.data, gamma, minf, pref, out.ctypes.data)
    return out


def physical_flux(w, n, gamma=1.4, pref=0.0) -> np.ndarray:
    out = np.empty(5)
    w, n = _a(w), _a(n)
    lib().oracle_physical_flux(w.ctypes.data, n.ctypes.data, gamma, pref, out.ctypes.data)
    return out


def boundary_value(wg, kind, data, n, gamma=1.4, rgas=287.0, pref=0.0) -> np.ndarray:
    out = np.empty(5)
    wg, data, n = _a(wg), _a(data), _a(n)
    lib().oracle_boundary_value(wg.ctypes.data, float(kind), data.ctypes.data, n.ctypes.data, gamma, rgas, pref,
                                out.ctypes.data)
    return out


def to_primitive(dq, ur, gamma=1.4) -> np.ndarray:
    out = np.empty(5)
    dq, ur = _a(dq), _a(ur)
    lib().oracle_to_primitive(dq.ctypes.data, ur.ctypes.data, gamma, out.ctypes.data)
    return out


def face_jacobian(kind: int, wl, wr, n, gamma=1.4, minf=0.5, pref=0.0, policy=0) -> np.ndarray:
    """(2, 5, 5): dF/dW_L and dF/dW_R of one face's flux (not area-weighted), forward-mode through
    the flux body in the policy's compute type."""
    out = np.empty(50)
    wl, wr, n = _a(wl), _a(wr), _a(n)
    lib().oracle_face_jacobian(policy, kind, wl.ctypes.data, wr.ctypes.data, n.ctypes.data, gamma, minf, pref,
                               out.ctypes.data)
    return out.reshape(2, 5, 5)


def boundary_jacobian(kind: int, wf, bc_kind, data, n, gamma=1.4, rgas=287.0, minf=0.5, pref=0.0, policy=0):
    """(5, 5): dF/dW_f of one boundary face, through boundary_value (the pressure flux's for slip
    and wall faces)."""
    out = np.empty(25)
    wf, data, n = _a(wf), _a(data), _a(n)
    lib().oracle_boundary_jacobian(policy, kind, wf.ctypes.data, float(bc_kind), data.ctypes.data, n.ctypes.data,
                                   gamma, rgas, minf, pref, out.ctypes.data)
    return out.reshape(5, 5)


def dwdu(dq, ur, gamma=1.4, policy=0) -> np.ndarray:
    """(5, 5): dW/dq of one cell's primitives, q the stored state."""
    out = np.empty(25)
    dq, ur = _a(dq), _a(ur)
    lib().oracle_dwdu(policy, dq.ctypes.data, ur.ctypes.data, gamma, out.ctypes.data)
    return out.reshape(5, 5)


def venkat(d, dmax, dmin, eps2) -> float:
    return lib().oracle_venkat(d, dmax, dmin, eps2)


def warp(w, gl, pl, gr, pr, m, gamma=1.4, pref=0.0) -> np.ndarray:
    out = np.empty(5)
    args = [_a(v) for v in (w, gl, pl, gr, pr, m)]
    lib().oracle_warp(*[a.ctypes.data for a in args], gamma, pref, out.ctypes.data)
    return out


class Oracle:
    """The reference's host residual path on one mesh; arrays FP64 SoA (5 x n_cells)."""

    def __init__(self, mesh, scheme, policy="fp64", bc_kind=None, bc=None, source=None) -> None:
        self.n = mesh.n_cells
        self.nf = mesh.n_faces
        self.nb = mesh.n_boundary
        planar = mesh.all_planar if scheme.planar is None else scheme.planar
        s = _Scheme()
        s.flux = {"roe": 0, "hllc": 1, "ausm+up": 2}[scheme.flux] if isinstance(scheme.flux, str) else scheme.flux
        s.second_order = int(scheme.second_order)
        s.limiter = int(scheme.limiter)
        s.planar = int(planar)
        s.gamma, s.gas_constant, s.mach_ref = scheme.gamma, scheme.gas_constant, scheme.mach_ref
        s.venkat_k, s.length = scheme.venkat_k, scheme.length
        s.scale[:] = [float(v) for v in scheme.scale]
        s.reference[:] = [float(v) for v in scheme.reference]
        keep = []

        def a(x, dtype=None):
            x = np.ascontiguousarray(x) if dtype is None else np.ascontiguousarray(x, dtype=dtype)
            keep.append(x)
            return x.ctypes.data

        nb = self.nb
        kind = np.zeros(nb) if bc_kind is None else bc_kind
        data = np.zeros(5 * nb) if bc is None else np.asarray(bc).reshape(-1)
        inp = _Inputs(
            mesh.n_cells, mesh.n_faces, mesh.n_interior,
            a(mesh.owner), a(mesh.neighbour) if mesh.n_interior else a(np.zeros(1, np.int32)),
            a(mesh.cell_face_offsets), a(mesh.cell_faces),
            a(mesh.face_normal), a(mesh.face_area), a(mesh.r_left), a(mesh.r_right), a(mesh.face_moment),
            a(mesh.cell_volume), a(mesh.stencil_offsets), a(mesh.stencil) if mesh.n_stencil else a(np.zeros(1, np.int32)),
            a(mesh.stencil_weights) if mesh.n_stencil else a(np.zeros(3)),
            a(kind, np.float64) if nb else a(np.zeros(1)), a(data, np.float64) if nb else a(np.zeros(5)),
            None if source is None else a(np.asarray(source).reshape(-1), np.float64),
        )  # fmt: skip
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        pol = {"fp64": 0, "fp32+acc64": 1}[policy] if isinstance(policy, str) else int(policy)
        if lib().oracle_create(pol, ctypes.byref(s), ctypes.byref(inp), ctypes.byref(h), err, 1024) != 0:
            raise RuntimeError(err.value.decode())
        self.h = h

    def __del__(self) -> None:
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def residual(self, u, gather: bool = True) -> np.ndarray:
        u = _a(u).reshape(-1)
        res = np.empty(5 * self.n)
        lib().oracle_residual(self.h, u.ctypes.data, res.ctypes.data, int(gather))
        return res

    def spectral_radius(self) -> np.ndarray:
        lam = np.empty(self.n)
        lib().oracle_spectral_radius(self.h, lam.ctypes.data)
        return lam

    def face_fluxes(self) -> np.ndarray:
        f = np.empty(5 * self.nf)
        lib().oracle_face_fluxes(self.h, f.ctypes.data)
        return f

    def arrays(self) -> dict[str, np.ndarray]:
        n = self.n
        out = {k: np.empty(m * n) for k, m in (("prim", 5), ("grad", 15), ("phi", 5), ("wmin", 5), ("wmax", 5))}
        out["prim_b_all"] = np.empty(5 * self.nf)
        lib().oracle_arrays(self.h, *[out[k].ctypes.data for k in ("prim", "grad", "phi", "wmin", "wmax", "prim_b_all")])
        pb = out.pop("prim_b_all").reshape(5, self.nf)[:, self.nf - self.nb :]
        out["prim_b"] = np.ascontiguousarray(pb).reshape(-1)
        return out

    def set_second_order(self, on: bool) -> None:
        lib().oracle_set_second_order(self.h, int(on))

    def face_force(self, u, faces) -> np.ndarray:
        """Euler::boundary_pressure's force on the listed boundary faces after the residual of u,
        shape (3, n)."""
        u = _a(u).reshape(-1)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        out = np.empty(3 * max(f.size, 1))
        lib().oracle_face_force(self.h, u.ctypes.data, f.ctypes.data, f.size, out.ctypes.data)
        return out[: 3 * f.size].reshape(3, f.size)

    def conservation(self, u) -> dict[str, np.ndarray]:
        """check_conservation's exact sums: E, D, boundary and faces, 5 each."""
        u = _a(u).reshape(-1)
        out = np.empty(20)
        lib().oracle_conservation(self.h, u.ctypes.data, out.ctypes.data)
        return {"E": out[0:5], "D": out[5:10], "boundary": out[10:15], "faces": out[15:20]}

    def first_invalid(self, u, res=None) -> int | None:
        u = _a(u).reshape(-1)
        r = None if res is None else _a(res).reshape(-1)
        c = lib().oracle_first_invalid(self.h, u.ctypes.data, None if r is None else r.ctypes.data)
        return None if c < 0 else c

    def residual_and_jacobian(self, u) -> dict[str, np.ndarray]:
        """Euler::residual_and_jacobian: the residual (gather loops) and the approximate Jacobian's
        parts in the device's layout (dwdu 25 x n, face_jac 50 x n_interior, bnd_jac 25 x
        n_boundary, entry (i, j) of a block at 5 i + j), FP64."""
        u = _a(u).reshape(-1)
        ni = self.nf - self.nb
        out = {"res": np.empty(5 * self.n), "dwdu": np.empty(25 * self.n), "face_jac": np.empty(50 * max(ni, 1)),
               "bnd_jac": np.empty(25 * max(self.nb, 1))}
        ptrs = [out[k].ctypes.data for k in ("res", "dwdu", "face_jac", "bnd_jac")]
        lib().oracle_residual_and_jacobian(self.h, u.ctypes.data, *ptrs)
        out["face_jac"] = out["face_jac"][: 50 * ni]
        out["bnd_jac"] = out["bnd_jac"][: 25 * self.nb]
        return out

    def step(self, u, cfl=0.0, dt=None, local=False, gather=True):
        """One SSP-RK3 step of u (copied): (new u, norms, time step, invalid cell)."""
        u = _a(u).reshape(-1).copy()
        r = np.empty(5)
        h = ctypes.c_double()
        inv = ctypes.c_int()
        lib().oracle_step(self.h, u.ctypes.data, cfl, int(dt is not None), 0.0 if dt is None else dt, int(local),
                          int(gather), r.ctypes.data, ctypes.byref(h), ctypes.byref(inv))
        return u, r, h.value, (None if inv.value < 0 else inv.value)


def implicit_solve(mesh, parts: dict, lam, cfl: float, gamma: float, u, oracle=None, second: bool = False):
    """One implicit solve and update on the Jacobian's parts (the device's layout), the spectral
    radius lam and the residual parts["res"] at state u: (lam / cfl I + dR/dU) du = -R by GCR (at
    most 10 directions, to 0.1) preconditioned by 4 symmetric multicolour block Gauss-Seidel sweeps,
    then u + alpha du with the admissible alpha. With `oracle` (the Oracle whose
    residual_and_jacobian made the parts) and `second` (they were made at second order), the Krylov
    operator adds the boundary faces' gradient term. Returns (new u, du, report, colours); report
    has cycles, residual_drop, alpha (0: no update), failed, singular."""
    n, ni = mesh.n_cells, mesh.n_interior
    u = _a(u).reshape(-1).copy()
    du = np.empty(5 * n)
    out = np.zeros(5)
    colours = np.empty(n, dtype=np.int32)
    keep = [
        np.ascontiguousarray(mesh.owner, dtype=np.int32),
        np.ascontiguousarray(mesh.neighbour, dtype=np.int32) if ni else np.zeros(1, np.int32),
        np.ascontiguousarray(mesh.cell_face_offsets, dtype=np.int64),
        np.ascontiguousarray(mesh.cell_faces, dtype=np.int32),
        _a(parts["dwdu"]),
        _a(parts["face_jac"]) if ni else np.zeros(1),
        _a(parts["bnd_jac"]) if mesh.n_boundary else np.zeros(1),
        _a(lam),
    ]
    res = _a(parts["res"]).reshape(-1)
    h = None if oracle is None else oracle.h
    lib().oracle_implicit_solve(h, int(bool(second)), n, ni, *[k.ctypes.data for k in keep], float(cfl), res.ctypes.data,
                                float(gamma),
                                u.ctypes.data, du.ctypes.data, out.ctypes.data, colours.ctypes.data)
    report = {"cycles": int(out[0]), "residual_drop": float(out[1]), "alpha": float(out[2]),
              "failed": bool(out[3]), "singular": bool(out[4])}
    return u, du, report, colours


def admissible_step(u, du, gamma: float) -> float:
    """EulerPhysics::admissible_step: the largest fraction (at most 1, halved down to 1e-6) of du
    that keeps every cell's density and pressure above half their values."""
    u, du = _a(u).reshape(-1), _a(du).reshape(-1)
    return lib().oracle_admissible_step(u.size // 5, u.ctypes.data, du.ctypes.data, float(gamma))
