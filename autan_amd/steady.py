"""Steady runs: the first-order start, local time steps and the convergence rule of the reference's
runtime (physics/src/euler_physics.cpp:1495-1571; runtime/src/run.cpp:2262-2389), run by the native
library (mxsteady_run) over the device stepper or over any object with step / evaluate /
set_second_order; and the steady implicit run (EulerPhysics::steady_step, pseudo-transient
continuation with switched evolution relaxation, GCR and block Gauss-Seidel; mximplicit_run) over
the device stepper or over any object with evaluate / set_second_order / jacobian / solve."""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from autan_amd import _lib

CONVERGED, BUDGET, DIVERGED = 0, 1, 2
END_NAMES = {0: "none", 1: "drop", 2: "stall", 3: "cap", 4: "level"}
STEADY_CFL = 0.8  # the device path's steady CFL number (physics/src/euler_physics.cpp:94)

_D5 = ctypes.c_double * 5
_STEP = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_int32))
_EVAL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32))
_ORDER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int32)


class _Ops(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("step", _STEP), ("evaluate", _EVAL), ("set_second_order", _ORDER)]


_SOLVE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double))
_CELL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32))


class _ImplicitOps(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("evaluate", _EVAL), ("set_second_order", _ORDER), ("jacobian", _EVAL),
                ("solve", _SOLVE), ("worst_cell", _CELL)]


class _Options(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int64),
        ("residual_orders", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("level_rule", ctypes.c_int32),
        ("start_order", ctypes.c_int32),
        ("scales", ctypes.c_void_p),
        ("reference", ctypes.c_void_p),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("first_order_end", ctypes.c_int32),
        ("first_order_steps", ctypes.c_int32),
        ("diverged_cell", ctypes.c_int32),
        ("plateau_mask", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("iterations", ctypes.c_int64),
        ("reference", ctypes.c_double * 5),
        ("residuals", ctypes.c_double * 5),
    ]


@dataclass
class SteadyResult:
    status: int
    iterations: int
    first_order_steps: int
    first_order_end: str
    diverged_cell: int | None
    plateau_mask: int
    reference: np.ndarray
    residuals: np.ndarray
    history: np.ndarray  # (iterations, 5) reported residuals
    provisional: np.ndarray  # (iterations,) bool

    @property
    def converged(self) -> bool:
        return self.status == CONVERGED


def _options(max_iterations, residual_orders, cfl, level_rule, start_order, scales, reference, keep) -> _Options:
    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != 5:
            raise ValueError("scales and reference hold 5 values")
        keep.append(a)
        return a.ctypes.data

    if start_order not in ("first", "second"):
        raise ValueError("start_order is 'first' or 'second'")
    return _Options(int(max_iterations), float(residual_orders), float(cfl), int(bool(level_rule)),
                    1 if start_order == "first" else 2, ptr(scales), ptr(reference))


def _finish(res: _Result, hist: np.ndarray, prov: np.ndarray) -> SteadyResult:
    n = int(res.iterations)
    return SteadyResult(
        status=int(res.status),
        iterations=n,
        first_order_steps=int(res.first_order_steps),
        first_order_end=END_NAMES[int(res.first_order_end)],
        diverged_cell=None if res.diverged_cell < 0 else int(res.diverged_cell),
        plateau_mask=int(res.plateau_mask),
        reference=np.array(res.reference[:]),
        residuals=np.array(res.residuals[:]),
        history=hist.reshape(-1, 5)[:n].copy(),
        provisional=prov[:n].astype(bool),
    )


def solve(stepper, max_iterations: int = 10000, residual_orders: float = 8.0, cfl: float = STEADY_CFL,
          level_rule: bool | None = None, start_order: str = "first", scales=None, reference=None) -> SteadyResult:
    """A steady run of the device stepper (autan_amd.stepper.DeviceStepper) from its current state.
    level_rule defaults to the stepper's policy (fp32+acc64); scales (Phi_k) enable the fp32+acc64
    round-off floors."""
    lib = _lib.euler()
    if level_rule is None:
        level_rule = stepper.policy == 1
    keep: list = []
    opts = _options(max_iterations, residual_orders, cfl, level_rule, start_order, scales, reference, keep)
    res = _Result()
    hist = np.zeros(5 * max_iterations)
    prov = np.zeros(max_iterations, dtype=np.int32)
    err = ctypes.create_string_buffer(2048)
    rc = lib.mxeuler_steady(stepper.handle, ctypes.byref(opts), ctypes.byref(res), ctypes.c_void_p(hist.ctypes.data),
                            ctypes.c_void_p(prov.ctypes.data), err, ctypes.c_size_t(len(err)))
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return _finish(res, hist, prov)


def solve_with(target, max_iterations: int = 10000, residual_orders: float = 8.0, cfl: float = STEADY_CFL,
               level_rule: bool = False, start_order: str = "first", scales=None, reference=None) -> SteadyResult:
    """The same run over a Python object with step(cfl) -> (residuals[5], invalid or None),
    evaluate() -> (residuals[5], invalid or None) and set_second_order(on)."""
    lib = _lib.euler()
    failure: list[BaseException] = []

    def guard(fn):
        def wrapped(*args):
            try:
                return fn(*args)
            except BaseException as e:  # reported after the native call returns
                failure.append(e)
                return 1

        return wrapped

    @guard
    def step(_ctx, c, out, inv):
        r, bad = target.step(c)
        for k in range(5):
            out[k] = float(r[k])
        inv[0] = -1 if bad is None else int(bad)
        return 0

    @guard
    def evaluate(_ctx, out, inv):
        r, bad = target.evaluate()
        for k in range(5):
            out[k] = float(r[k])
        inv[0] = -1 if bad is None else int(bad)
        return 0

    @guard
    def order(_ctx, on):
        target.set_second_order(bool(on))
        return 0

    ops = _Ops(None, _STEP(step), _EVAL(evaluate), _ORDER(order))
    keep: list = []
    opts = _options(max_iterations, residual_orders, cfl, level_rule, start_order, scales, reference, keep)
    res = _Result()
    hist = np.zeros(5 * max_iterations)
    prov = np.zeros(max_iterations, dtype=np.int32)
    err = ctypes.create_string_buffer(2048)
    rc = lib.mxsteady_run(ctypes.byref(ops), ctypes.byref(opts), ctypes.byref(res), ctypes.c_void_p(hist.ctypes.data),
                          ctypes.c_void_p(prov.ctypes.data), err, ctypes.c_size_t(len(err)))
    if failure:
        raise failure[0]
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return _finish(res, hist, prov)


@dataclass
class ImplicitResult(SteadyResult):
    cfl: np.ndarray = None  # per iteration: the CFL number
    cycles: np.ndarray = None  # the GCR directions taken
    linear_drop: np.ndarray = None  # the solve's residual drop
    alpha: np.ndarray = None  # the update fraction (0: rejected)


def _implicit_finish(res, hist, prov, stats) -> ImplicitResult:
    base = _finish(res, hist, prov)
    n = base.iterations
    st = stats.reshape(-1, 4)[:n]
    return ImplicitResult(**base.__dict__, cfl=st[:, 0].copy(), cycles=st[:, 1].astype(int), linear_drop=st[:, 2].copy(),
                          alpha=st[:, 3].copy())


def solve_implicit(stepper, max_iterations: int = 500, residual_orders: float = 8.0, level_rule: bool | None = None,
                   start_order: str = "first", scales=None, reference=None) -> ImplicitResult:
    """The steady implicit run of the device stepper (one case) from its current state: per
    iteration R(u) and the Jacobian on the device, the CFL number from the relative residual, one
    solve and update (mxeuler_implicit_steady). level_rule defaults to the stepper's policy."""
    lib = _lib.euler()
    if level_rule is None:
        level_rule = stepper.policy == 1
    keep: list = []
    opts = _options(max_iterations, residual_orders, STEADY_CFL, level_rule, start_order, scales, reference, keep)
    res = _Result()
    hist = np.zeros(5 * max_iterations)
    prov = np.zeros(max_iterations, dtype=np.int32)
    stats = np.zeros(4 * max_iterations)
    err = ctypes.create_string_buffer(2048)
    rc = lib.mxeuler_implicit_steady(stepper.handle, ctypes.byref(opts), ctypes.byref(res),
                                     ctypes.c_void_p(hist.ctypes.data), ctypes.c_void_p(prov.ctypes.data),
                                     ctypes.c_void_p(stats.ctypes.data), err, ctypes.c_size_t(len(err)))
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return _implicit_finish(res, hist, prov, stats)


def solve_implicit_with(target, max_iterations: int = 500, residual_orders: float = 8.0, level_rule: bool = False,
                        start_order: str = "first", scales=None, reference=None) -> ImplicitResult:
    """The same run (mximplicit_run) over a Python object with evaluate() -> (residuals[5], invalid
    or None), set_second_order(on), jacobian() -> (residuals[5], invalid or None), solve(cfl) ->
    report (5 values: GCR directions, linear drop, alpha, failed, singular) and optionally
    worst_cell() -> int."""
    lib = _lib.euler()
    failure: list[BaseException] = []

    def guard(fn):
        def wrapped(*args):
            try:
                return fn(*args)
            except BaseException as e:  # reported after the native call returns
                failure.append(e)
                return 1

        return wrapped

    def norms_call(method):
        @guard
        def call(_ctx, out, inv):
            r, bad = method()
            for k in range(5):
                out[k] = float(r[k])
            inv[0] = -1 if bad is None else int(bad)
            return 0

        return call

    @guard
    def order(_ctx, on):
        target.set_second_order(bool(on))
        return 0

    @guard
    def solve(_ctx, cfl, out):
        rep = target.solve(cfl)
        for k in range(5):
            out[k] = float(rep[k])
        return 0

    @guard
    def worst(_ctx, cell):
        cell[0] = int(target.worst_cell())
        return 0

    ops = _ImplicitOps(None, _EVAL(norms_call(target.evaluate)), _ORDER(order), _EVAL(norms_call(target.jacobian)),
                       _SOLVE(solve), _CELL(worst) if hasattr(target, "worst_cell") else _CELL())
    keep: list = []
    opts = _options(max_iterations, residual_orders, STEADY_CFL, level_rule, start_order, scales, reference, keep)
    res = _Result()
    hist = np.zeros(5 * max_iterations)
    prov = np.zeros(max_iterations, dtype=np.int32)
    stats = np.zeros(4 * max_iterations)
    err = ctypes.create_string_buffer(2048)
    rc = lib.mximplicit_run(ctypes.byref(ops), ctypes.byref(opts), ctypes.byref(res), ctypes.c_void_p(hist.ctypes.data),
                            ctypes.c_void_p(prov.ctypes.data), ctypes.c_void_p(stats.ctypes.data), err,
                            ctypes.c_size_t(len(err)))
    if failure:
        raise failure[0]
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return _implicit_finish(res, hist, prov, stats)
