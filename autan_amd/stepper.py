"""Python mirror of the reference's DeviceStepper<P> over the HIP library (include/mxeuler.h).

`DeviceStepper(mesh, scheme, policy, ...)` is make_device_stepper<P>
(physics/include/autan/physics/euler_device.hpp:49-54); its methods are the virtual interface of
physics/include/autan/physics/euler_device.hpp:22-41: step, evaluate, residual, set_second_order,
upload, download. Errors raise RuntimeError carrying the library's message (E1101 for a device
failure). There is no CPU fallback: without the built library every call fails.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from autan_amd import _lib
from autan_amd.mesh import Mesh

FP64, FP32_ACC64 = 0, 1
ROE, HLLC, AUSM_UP = 0, 1, 2
FLUXES = {"roe": ROE, "hllc": HLLC, "ausm+up": AUSM_UP}
POLICIES = {"fp64": FP64, "fp32+acc64": FP32_ACC64}

_d = ctypes.c_double
_i32 = ctypes.c_int32
_p = ctypes.c_void_p


class _Scheme(ctypes.Structure):
    _fields_ = [
        ("flux", _i32),
        ("second_order", _i32),
        ("limiter", _i32),
        ("planar", _i32),
        ("gamma", _d),
        ("gas_constant", _d),
        ("mach_ref", _d),
        ("venkat_k", _d),
        ("length", _d),
        ("scale", _d * 5),
        ("reference", _d * 5),
    ]


class _Topology(ctypes.Structure):
    _fields_ = [
        ("n_cells", _i32),
        ("n_faces", _i32),
        ("n_interior", _i32),
        ("owner", _p),
        ("neighbour", _p),
        ("cell_face_offsets", _p),
        ("cell_faces", _p),
    ]


class _Geometry(ctypes.Structure):
    _fields_ = [
        ("face_normal", _p),
        ("face_area", _p),
        ("r_left", _p),
        ("r_right", _p),
        ("face_moment", _p),
        ("cell_volume", _p),
        ("all_planar", _i32),
    ]


class _Stencil(ctypes.Structure):
    _fields_ = [("offsets", _p), ("entries", _p), ("weights", _p)]


@dataclass
class Scheme:
    """Scheme and Gas of physics/include/autan/physics/euler.hpp:39-71."""

    flux: str = "roe"
    second_order: bool = True
    limiter: bool = True
    planar: bool | None = None  # None: the planar loops wherever every face is proven planar
    gamma: float = 1.4
    gas_constant: float = 287.0
    mach_ref: float = 0.3
    venkat_k: float = 10.0
    length: float = 1.0
    scale: tuple = (1.0, 1.0, 1.0, 1.0, 1.0)
    reference: tuple = (1.0, 0.0, 0.0, 0.0, 2.5)

    def native(self, all_planar: bool) -> _Scheme:
        planar = all_planar if self.planar is None else self.planar
        s = _Scheme()
        s.flux = FLUXES[self.flux] if isinstance(self.flux, str) else int(self.flux)
        s.second_order = int(self.second_order)
        s.limiter = int(self.limiter)
        s.planar = int(planar)
        s.gamma, s.gas_constant, s.mach_ref = self.gamma, self.gas_constant, self.mach_ref
        s.venkat_k, s.length = self.venkat_k, self.length
        s.scale[:] = [float(v) for v in self.scale]
        s.reference[:] = [float(v) for v in self.reference]
        return s


def freestream_scheme(rho: float, velocity, p: float, gamma: float = 1.4, **kw) -> Scheme:
    """EulerPhysics::scheme (physics/src/euler_physics.cpp:594-616): q scales and U_ref from the
    freestream, M_co = |v| / c."""
    v = np.asarray(velocity, dtype=float)
    speed = float(np.sqrt(v @ v))
    c = float(np.sqrt(gamma * p / rho))
    s = max(speed, c)
    q2 = float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    ref = (rho, rho * v[0], rho * v[1], rho * v[2], p / (gamma - 1.0) + 0.5 * rho * q2)
    kw.setdefault("mach_ref", speed / c if speed > 0 else 0.3)
    return Scheme(gamma=gamma, scale=(rho, s, s, s, p), reference=ref, **kw)


@dataclass
class Result:
    """DeviceStepper::Result: RMS residual norms of R(u), the time step, the invalid cell."""

    residuals: np.ndarray
    time_step: float
    invalid: int | None


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def _err() -> ctypes.Array:
    return ctypes.create_string_buffer(2048)


def _check(rc: int, err) -> None:
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))


# MXEULER_STATIC_FORCE_FACES: the most faces one mxeuler_face_force_noalloc call stages.
STATIC_FORCE_FACES = 131072


class DeviceStepper:
    """make_device_stepper<P> over the HIP library; arrays are FP64 SoA (5 x n_cells)."""

    def __init__(
        self,
        mesh: Mesh,
        scheme: Scheme,
        policy: str | int = "fp32+acc64",
        bc_kind: np.ndarray | None = None,
        bc: np.ndarray | None = None,
        source: np.ndarray | None = None,
        stream: int | None = None,
    ) -> None:
        self.lib = _lib.euler()
        self.mesh = mesh
        self.n = mesh.n_cells
        self.policy = POLICIES[policy] if isinstance(policy, str) else int(policy)
        nb = mesh.n_boundary
        self._keep = []  # arrays whose addresses the library reads during create
        kind = np.zeros(nb) if bc_kind is None else np.ascontiguousarray(bc_kind, dtype=np.float64)
        data = np.zeros(5 * nb) if bc is None else np.ascontiguousarray(bc, dtype=np.float64).reshape(-1)
        if kind.size != nb or data.size != 5 * nb:
            raise ValueError("boundary arrays do not match the boundary faces")
        src = None if source is None else np.ascontiguousarray(source, dtype=np.float64).reshape(-1)
        if src is not None and src.size != 5 * self.n:
            raise ValueError("source array does not match the cells")
        sc = scheme.native(mesh.all_planar)
        topo = _Topology(
            mesh.n_cells,
            mesh.n_faces,
            mesh.n_interior,
            self._a(mesh.owner),
            self._a(mesh.neighbour),
            self._a(mesh.cell_face_offsets),
            self._a(mesh.cell_faces),
        )
        geo = _Geometry(
            self._a(mesh.face_normal),
            self._a(mesh.face_area),
            self._a(mesh.r_left),
            self._a(mesh.r_right),
            self._a(mesh.face_moment),
            self._a(mesh.cell_volume),
            int(mesh.all_planar),
        )
        st = _Stencil(self._a(mesh.stencil_offsets), self._a(mesh.stencil), self._a(mesh.stencil_weights))
        handle = ctypes.c_void_p()
        err = _err()
        rc = self.lib.mxeuler_create(
            _i32(self.policy),
            ctypes.byref(sc),
            ctypes.byref(topo),
            ctypes.byref(geo),
            ctypes.byref(st),
            _p(_ptr(kind)),
            _p(_ptr(data)),
            _p(_ptr(src)),
            _p(stream),
            ctypes.byref(handle),
            err,
            ctypes.c_size_t(len(err)),
        )
        self._keep = []
        _check(rc, err)
        self.handle = handle

    def _a(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        self._keep.append(a)
        return a.ctypes.data

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.mxeuler_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        self.close()

    def __enter__(self) -> "DeviceStepper":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _call(self, fn, *args) -> None:
        err = _err()
        _check(fn(self.handle, *args, err, ctypes.c_size_t(len(err))), err)

    def _state(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 5 * self.n:
            raise ValueError("state array does not match the cells")
        return u

    def upload(self, u: np.ndarray) -> None:
        u = self._state(u)
        self._call(self.lib.mxeuler_upload, _p(u.ctypes.data))

    def download(self) -> np.ndarray:
        u = np.empty(5 * self.n)
        self._call(self.lib.mxeuler_download, _p(u.ctypes.data))
        return u

    def residual(self, u: np.ndarray) -> np.ndarray:
        u = self._state(u)
        res = np.empty(5 * self.n)
        self._call(self.lib.mxeuler_residual, _p(u.ctypes.data), _p(res.ctypes.data))
        return res

    def residual_and_jacobian(self, u: np.ndarray) -> dict[str, np.ndarray]:
        """Euler::residual_and_jacobian: R(u) (5, n) and the approximate Jacobian's parts in FP64,
        the reconstruction's gradients and limiter held fixed: dwdu (5, 5, n) dW/dq per cell,
        face_jac (2, 5, 5, n_interior) dF/dW_L and dF/dW_R per interior face (area-weighted),
        bnd_jac (5, 5, n_boundary) dF/dW_L per boundary face; on warped faces the face body's warp
        term's derivative is included."""
        u = self._state(u)
        n, ni, nb = self.n, self.mesh.n_interior, self.mesh.n_boundary
        res = np.empty(5 * n)
        dwdu = np.empty(25 * n)
        fj = np.empty(50 * max(ni, 1))
        bj = np.empty(25 * max(nb, 1))
        self._call(self.lib.mxeuler_residual_and_jacobian, _p(u.ctypes.data), _p(res.ctypes.data),
                   _p(dwdu.ctypes.data), _p(fj.ctypes.data), _p(bj.ctypes.data))
        return {"res": res.reshape(5, n), "dwdu": dwdu.reshape(5, 5, n),
                "face_jac": fj[: 50 * ni].reshape(2, 5, 5, ni), "bnd_jac": bj[: 25 * nb].reshape(5, 5, nb)}

    def implicit_evaluate(self) -> tuple[np.ndarray, int | None]:
        """The implicit step's evaluation of the stepper's state (mxeuler_implicit_evaluate): the
        norms of R(u) at the order set and the first invalid cell; the residual, spectral radius and
        Jacobian stay on the device for implicit_solve."""
        r = np.empty(5)
        inv = ctypes.c_int32()
        self._call(self.lib.mxeuler_implicit_evaluate, _p(r.ctypes.data), ctypes.byref(inv))
        return r, (None if inv.value < 0 else inv.value)

    def implicit_solve(self, cfl: float) -> dict:
        """One solve and update (mxeuler_implicit_solve) at this CFL number: GCR directions taken,
        the solve's residual drop, the update fraction alpha (0: none), failed, singular."""
        rep = np.zeros(5)
        self._call(self.lib.mxeuler_implicit_solve, ctypes.c_double(cfl), _p(rep.ctypes.data))
        return {"cycles": int(rep[0]), "residual_drop": float(rep[1]), "alpha": float(rep[2]), "failed": bool(rep[3]),
                "singular": bool(rep[4])}

    def implicit_colours(self) -> np.ndarray:
        """The colour of every cell in the implicit step's Gauss-Seidel sweeps."""
        col = np.empty(self.n, dtype=np.int32)
        self._call(self.lib.mxeuler_implicit_colours, _p(col.ctypes.data))
        return col

    def upload_device(self, u: int, stream: int = 0) -> None:
        """The state from device memory: `u` the address of 5 x n FP64 values on the stepper's GPU,
        allocated through the HIP runtime this library links (/opt/rocm's libamdhip64); the copy
        is ordered after the work `stream` (a raw hipStream_t of that runtime, 0: the null stream)
        holds, and `stream`'s later work after it."""
        self._call(self.lib.mxeuler_upload_device, _p(u), _p(stream))

    def download_device(self, u: int, stream: int = 0) -> None:
        """The state into device memory, ordered as upload_device."""
        self._call(self.lib.mxeuler_download_device, _p(u), _p(stream))

    def residual_device(self, u: int, res: int, stream: int = 0) -> None:
        """R(u) into res, both device addresses of 5 x n FP64 values, ordered as upload_device."""
        self._call(self.lib.mxeuler_residual_device, _p(u), _p(res), _p(stream))

    def step(self, cfl: float = 0.0, dt: float | None = None, local: bool = False) -> Result:
        r = (ctypes.c_double * 5)()
        h = ctypes.c_double()
        inv = ctypes.c_int32()
        self._call(
            self.lib.mxeuler_step,
            _d(cfl),
            _i32(dt is not None),
            _d(0.0 if dt is None else dt),
            _i32(local),
            r,
            ctypes.byref(h),
            ctypes.byref(inv),
        )
        return Result(np.array(r[:]), h.value, None if inv.value < 0 else inv.value)

    def evaluate(self) -> Result:
        r = (ctypes.c_double * 5)()
        inv = ctypes.c_int32()
        self._call(self.lib.mxeuler_evaluate, r, ctypes.byref(inv))
        return Result(np.array(r[:]), 0.0, None if inv.value < 0 else inv.value)

    def set_second_order(self, on: bool) -> None:
        self._call(self.lib.mxeuler_set_second_order, _i32(on))

    def set_boundary_data(self, bc: np.ndarray) -> None:
        bc = np.ascontiguousarray(bc, dtype=np.float64).reshape(-1)
        self._call(self.lib.mxeuler_set_boundary_data, _p(bc.ctypes.data))

    def set_source(self, source: np.ndarray | None) -> None:
        s = None if source is None else np.ascontiguousarray(source, dtype=np.float64).reshape(-1)
        self._call(self.lib.mxeuler_set_source, _p(_ptr(s)))

    def spectral_radius(self) -> np.ndarray:
        lam = np.empty(self.n)
        self._call(self.lib.mxeuler_spectral_radius, _p(lam.ctypes.data))
        return lam

    def arrays(self) -> dict[str, np.ndarray]:
        n, nb = self.n, self.mesh.n_boundary
        out = {
            "prim": np.empty(5 * n),
            "grad": np.empty(15 * n),
            "phi": np.empty(5 * n),
            "prim_b": np.empty(5 * max(nb, 1)),
        }
        self._call(self.lib.mxeuler_arrays, *[_p(out[k].ctypes.data) for k in ("prim", "grad", "phi", "prim_b")])
        out["prim_b"] = out["prim_b"][: 5 * nb]
        return out

    def face_force(self, faces: np.ndarray) -> np.ndarray:
        """Euler::boundary_pressure's pressure force on the listed boundary faces for the current
        state, shape (3, n)."""
        f = np.ascontiguousarray(faces, dtype=np.int32)
        out = np.empty(3 * max(f.size, 1))
        self._call(self.lib.mxeuler_face_force, _p(f.ctypes.data), ctypes.c_int64(f.size), _p(out.ctypes.data))
        return out[: 3 * f.size].reshape(3, f.size)

    def face_force_noalloc(self, faces: np.ndarray) -> np.ndarray:
        """face_force staged in the library's static device array (mxeuler_face_force_noalloc): no
        device allocation, at most STATIC_FORCE_FACES faces; the same results."""
        f = np.ascontiguousarray(faces, dtype=np.int32)
        out = np.empty(3 * max(f.size, 1))
        self._call(self.lib.mxeuler_face_force_noalloc, _p(f.ctypes.data), ctypes.c_int64(f.size),
                   _p(out.ctypes.data))
        return out[: 3 * f.size].reshape(3, f.size)

    def loads(self, faces: np.ndarray, center=(0.0, 0.0, 0.0), p_inf: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        """EulerPhysics::loads over the listed boundary faces: force and moment about `center`
        (exact sums), the pressure force minus p_inf S."""
        f = np.ascontiguousarray(faces, dtype=np.int32)
        nf = self.mesh.n_faces
        av = np.ascontiguousarray(self.mesh.face_area_vector.reshape(3, nf)[:, f])
        fc = np.ascontiguousarray(self.mesh.face_centroid.reshape(3, nf)[:, f])
        c = np.ascontiguousarray(center, dtype=np.float64)
        force = (ctypes.c_double * 3)()
        moment = (ctypes.c_double * 3)()
        self._call(
            self.lib.mxeuler_loads,
            _p(f.ctypes.data),
            ctypes.c_int64(f.size),
            _p(av.ctypes.data),
            _p(fc.ctypes.data),
            _p(c.ctypes.data),
            _d(p_inf),
            force,
            moment,
        )
        return np.array(force[:]), np.array(moment[:])

    def autotune(self, steps: int = 5, cfl: float = 0.5, local: bool = False) -> str:
        """Keeps the fastest kernel variant (same bits); the state is restored."""
        report = ctypes.create_string_buffer(2048)
        self._call(self.lib.mxeuler_autotune, _i32(steps), _d(cfl), _i32(local), report, ctypes.c_size_t(len(report)))
        return report.value.decode()

    def tune(self, key: str, value: int) -> None:
        """Selects a kernel variant (same bits, different speed): "face_stage" 0/1, "graph" 0/1,
        "grad_occupancy" 0/7/8, "face_occupancy" 0/5/8, "face_block" 256/512/1024, "grad_nontemporal"
        0/1 (default 1)."""
        self._call(self.lib.mxeuler_tune, ctypes.c_char_p(key.encode()), _i32(value))

    def conservation(self) -> dict[str, np.ndarray]:
        """EulerPhysics::check_conservation's exact sums on the current state: E = sum R^F - sum
        F_b, D = sum |F| over the incidences, the boundary fluxes and sum R^F, 5 each."""
        out = (ctypes.c_double * 20)()
        self._call(self.lib.mxeuler_conservation, out)
        a = np.array(out[:])
        return {"E": a[0:5], "D": a[5:10], "boundary": a[10:15], "faces": a[15:20]}

    def residual_scales(self, density: float, speed: float, pressure: float) -> np.ndarray:
        """EulerPhysics::residual_scales: Phi_k = f_k sqrt(sum A_i^2 / N)."""
        out = (ctypes.c_double * 5)()
        self._call(self.lib.mxeuler_residual_scales, _d(density), _d(speed), _d(pressure), out)
        return np.array(out[:])

    def profile(self, steps: int, cfl: float, local: bool = False, max_kernels: int = 32) -> dict:
        """Milliseconds per step and per launch of each kernel over `steps` steps (HIP events)."""
        ms_step = ctypes.c_double()
        ms = (ctypes.c_double * max_kernels)()
        launches = (ctypes.c_int32 * max_kernels)()
        nk = ctypes.c_int32()
        names = ctypes.create_string_buffer(4096)
        self._call(
            self.lib.mxeuler_profile,
            _i32(steps),
            _d(cfl),
            _i32(local),
            ctypes.byref(ms_step),
            ms,
            launches,
            _i32(max_kernels),
            ctypes.byref(nk),
            names,
            ctypes.c_size_t(len(names)),
        )
        keys = names.value.decode().split(";")[: nk.value]
        return {
            "ms_per_step": ms_step.value,
            "kernels": {k: {"ms": ms[i], "launches": launches[i]} for i, k in enumerate(keys)},
        }


def probe_math(policy: int, x: np.ndarray, y: np.ndarray) -> dict[str, np.ndarray]:
    """sqrt, x / y, exp and log evaluated on the device (FP64 policy 0, FP32 policy 1)."""
    lib = _lib.euler()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = {k: np.empty(x.size) for k in ("sqrt", "div", "exp", "log")}
    err = _err()
    rc = lib.mxeuler_probe_math(
        _i32(policy),
        _p(x.ctypes.data),
        _p(y.ctypes.data),
        ctypes.c_int64(x.size),
        *[_p(out[k].ctypes.data) for k in ("sqrt", "div", "exp", "log")],
        err,
        ctypes.c_size_t(len(err)),
    )
    _check(rc, err)
    return out
