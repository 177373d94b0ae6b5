"""Observed order of accuracy from the manufactured solution (docs/design/euler.md, Manufactured
solution; physics/mms/euler_mms.py): per mesh level, the steady discrete solution of R(u) = 0 with the
exact state on every boundary face and the exact source at the cell centroids, and the
volume-weighted L2 error of each conserved variable at the centroids; the order is the least-squares
slope of log error against log h with h = N^(-1/3), and the finest pair's.

The reference reaches the steady solution by Newton-Krylov on its residual; here the device
stepper's steady run does (local time steps, SSP-RK3, second order from the first step, starting
from the exact state): the discrete solution is the same fixed point whichever iteration reaches it,
and the run reports how far its residual dropped. The study stops on the residual drop or the
iteration budget alone: the fp32+acc64 round-off plateau rule of a solver run ends a run once its
residual is within ten units of FP32 round-off of the flux scale, while smooth error modes, whose
residual is a mesh spacing times smaller than their error, are still decaying.
"""

from __future__ import annotations

import math
import time

import numpy as np

from autan_amd import steady, workloads
from autan_amd.stepper import DeviceStepper

CONSERVED = ("rho", "rhou", "rhov", "rhow", "rhoE")


def errors(volume: np.ndarray, u: np.ndarray, exact: np.ndarray) -> np.ndarray:
    """Volume-weighted L2 norm of u - exact per conserved variable: sqrt(sum V e^2 / sum V)."""
    v = np.asarray(volume, dtype=np.float64)
    e = (np.asarray(u, dtype=np.float64) - np.asarray(exact, dtype=np.float64)).reshape(5, -1)
    return np.sqrt((e * e) @ v / v.sum())


def fit_order(h, e) -> tuple[float, float]:
    """The least-squares slope of log e against log h, and the order of the finest pair (the two
    smallest h); at least two levels."""
    if len(h) < 2 or len(h) != len(e):
        raise ValueError("an order needs errors on two levels at least")
    lh, le = np.log(np.asarray(h, dtype=np.float64)), np.log(np.asarray(e, dtype=np.float64))
    slope = float(np.polyfit(lh, le, 1)[0])
    i = np.argsort(lh)[:2]
    pair = float((le[i[1]] - le[i[0]]) / (lh[i[1]] - lh[i[0]]))
    return slope, pair


def level(n: int, policy: str = "fp32+acc64", flux: str = "hllc", limiter: bool = True,
          max_iterations: int = 20000, residual_orders: float = 8.0, log=print, chunk: int = 4000) -> dict:
    """One level: the Kuhn tetrahedra n^3 x 6 of the benchmark, its steady discrete solution and its
    errors against the manufactured solution at the centroids."""
    w = workloads.tets(n, policy, flux)
    w.scheme.limiter = limiter
    m = w.mesh
    dev = DeviceStepper(m, w.scheme, w.policy, w.bc_kind, w.bc, w.source)
    dev.upload(w.u)
    # In chunks, each dropping from the first chunk's iteration-1 residuals, so a long run reports
    # its progress as it goes.
    t0 = time.perf_counter()
    run = steady.solve(dev, max_iterations=min(chunk, max_iterations), residual_orders=residual_orders,
                       start_order="second", level_rule=False)
    reference = run.reference
    iterations = run.iterations
    while run.status == steady.BUDGET and iterations < max_iterations:
        log(f"order: n={n} {policy}: {iterations} iterations, residual drop "
            f"{[round(-math.log10(a / b), 2) if a > 0 and b > 0 else None for a, b in zip(run.residuals, reference)]}")
        run = steady.solve(dev, max_iterations=min(chunk, max_iterations - iterations), residual_orders=residual_orders,
                           start_order="second", level_rule=False, reference=reference)
        iterations += run.iterations
    sec = time.perf_counter() - t0
    u = dev.download()
    vol = m.arrays["cell_volume"]
    e = errors(vol, u, w.u)
    drop = [float(a / b) if b > 0 and a > 0 else None for a, b in zip(run.residuals, reference)]
    log(f"order: n={n} {policy} {flux} limiter={limiter}: {iterations} iterations in {sec:.1f} s, "
        f"status {run.status}, errors {e}")
    return {
        "n": n,
        "cells": m.n_cells,
        "h": m.n_cells ** (-1.0 / 3.0),
        "policy": policy,
        "iterations": int(iterations),
        "status": ["converged", "budget", "diverged"][run.status],
        "seconds": round(sec, 2),
        "residual_drop_orders": [None if d is None else round(-math.log10(d), 2) for d in drop],
        "errors": {k: float(v) for k, v in zip(CONSERVED, e)},
        "state": u,
    }


def study(levels=(64, 96, 128), policies=("fp64", "fp32+acc64"), flux: str = "hllc", limiter: bool = True,
          max_iterations: int = 20000, residual_orders: float = 8.0, log=print) -> dict:
    """Every level in every policy; fitted and finest-pair orders per conserved variable and policy,
    and each fp32+acc64 level's distance from the fp64 solution on the same mesh."""
    out: dict = {"levels": {}, "orders": {}, "fp32_vs_fp64": {}}
    for pol in policies:
        rows = [level(n, pol, flux, limiter, max_iterations, residual_orders, log) for n in levels]
        out["levels"][pol] = rows
        if len(rows) >= 2:  # an order needs two levels at least
            out["orders"][pol] = {
                k: [round(x, 3) for x in fit_order([r["h"] for r in rows], [r["errors"][k] for r in rows])]
                for k in CONSERVED
            }
    if "fp64" in policies and "fp32+acc64" in policies:
        for r64, r32 in zip(out["levels"]["fp64"], out["levels"]["fp32+acc64"]):
            m = workloads.tets(r64["n"], "fp64", flux).mesh
            d = errors(m.arrays["cell_volume"], r32["state"], r64["state"])
            out["fp32_vs_fp64"][str(r64["n"])] = {k: float(v) for k, v in zip(CONSERVED, d)}
    for pol in out["levels"]:
        for r in out["levels"][pol]:
            r.pop("state")
    return out
