"""The bump-channel parameter sweep: cases of uneven size, bump height, Mach number and flow angle
(fixed seed), stepped together to their steady state by one batch stepper.

Case k: cell count log-uniform in [cells_lo, cells_hi], bump height uniform in [h_lo, h_hi], Mach
uniform in [mach_lo, mach_hi], flow angle uniform in [alpha_lo, alpha_hi] degrees, drawn in that
order per case from numpy's default generator with the given seed. Each case is the hybrid bump
channel (prism rows on the wall, a band of merged polyhedra) on a lattice of cubic cells
3 n x n x n with n chosen for the drawn count; second-order HLLC (or the given flux) in
fp32+acc64, local time steps at CFL 0.8, to a 1e-8 relative drop or the round-off plateau rule.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from autan_amd import batch, workloads


@dataclass
class Spec:
    cells: int  # the drawn count
    h: float
    mach: float
    alpha: float
    lattice: tuple  # nx, ny, nz


def lattice_for(cells: float) -> tuple[int, int, int]:
    """A 3n x n x n lattice of about `cells` hexahedra (n >= 8, nx and nz even)."""
    n = max(8, int(round((cells / 3.0) ** (1.0 / 3.0))))
    nx = 3 * n + (3 * n) % 2
    nz = n + n % 2
    return nx, n, nz


def specs(n_cases: int = 64, seed: int = 1, cells=(2e4, 5e5), h=(0.02, 0.12), mach=(0.3, 0.8),
          alpha=(-4.0, 4.0)) -> list[Spec]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        n = math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))
        hh = rng.uniform(*h)
        m = rng.uniform(*mach)
        a = rng.uniform(*alpha)
        out.append(Spec(int(round(n)), float(hh), float(m), float(a), lattice_for(n)))
    return out


def build(spec: Spec, policy: str = "fp32+acc64", flux: str = "hllc") -> batch.Case:
    nx, ny, nz = spec.lattice
    w = workloads.bump(nx, ny, nz, policy, flux, h=spec.h, mach=spec.mach, alpha_deg=spec.alpha, prism_layers=2)
    w.scheme.planar = False
    return batch.Case(w.mesh, w.scheme, w.u, w.bc_kind, w.bc)


def residual_scales(case: batch.Case) -> np.ndarray:
    """Phi_k of EulerPhysics::residual_scales for the case's freestream (rho q A, rho q^2 A, rho H q A
    with A = sqrt(sum A_i^2 / N))."""
    m = case.mesh
    area = np.zeros(m.n_cells)
    np.add.at(area, m.owner, m.face_area)
    np.add.at(area, m.neighbour, m.face_area[: m.n_interior])
    a = math.sqrt(float(np.sum(area * area)) / m.n_cells)
    ref = np.array(case.scheme.reference)
    g = case.scheme.gamma
    speed = float(np.linalg.norm(ref[1:4])) / ref[0]
    p = (g - 1.0) * (ref[4] - 0.5 * ref[0] * speed * speed)
    c = math.sqrt(g * p / ref[0])
    q = speed + c
    hh = c * c / (g - 1.0) + 0.5 * speed * speed
    mass = ref[0] * q * a
    return np.array([mass, mass * q, mass * q, mass * q, mass * hh])


def run(n_cases: int = 64, seed: int = 1, max_iterations: int = 20000, residual_orders: float = 8.0,
        policy: str = "fp32+acc64", flux: str = "hllc", start_order: str = "first", autotune: bool = True,
        cells=(2e4, 5e5), log=print) -> dict:
    """The sweep: build, tune, solve; aggregate cells x iterations per second and per-case counts."""
    t0 = time.time()
    sp = specs(n_cases, seed, cells)
    cases = [build(s, policy, flux) for s in sp]
    log(f"sweep: {n_cases} cases, {sum(c.mesh.n_cells for c in cases)} cells, built in {time.time() - t0:.1f} s")
    b = batch.BatchStepper(cases, policy)
    tuned = b.autotune(steps=5) if autotune else None
    if tuned is not None:
        log(f"sweep: autotuned {tuned}")
    else:
        b.tune("graph", 1)
    b.upload()
    # One scale set for the round-off floors: each case's own.
    scales = np.array([residual_scales(c) for c in cases])
    t1 = time.perf_counter()
    runs = b.steady(max_iterations=max_iterations, residual_orders=residual_orders, start_order=start_order,
                    scales=scales, history=False)
    wall = time.perf_counter() - t1
    states = b.download()
    cells_iter = sum(c.mesh.n_cells * r.iterations for c, r in zip(cases, runs))
    status = ["converged", "budget", "diverged"]
    return {
        "cases": [
            {
                "cells": c.mesh.n_cells,
                "h": s.h,
                "mach": s.mach,
                "alpha": s.alpha,
                "iterations": r.iterations,
                "status": status[r.status],
                "first_order_steps": r.first_order_steps,
                "plateau_mask": r.plateau_mask,
                "orders": [round(math.log10(a / bb), 2) if a > 0 and bb > 0 else None
                           for a, bb in zip(r.reference, r.residuals)],
            }
            for c, s, r in zip(cases, sp, runs)
        ],
        "wall_s": wall,
        "cells_iter_per_s": cells_iter / wall,
        "total_cells": sum(c.mesh.n_cells for c in cases),
        "tuned": tuned,
        "states": states,
        "specs": sp,
        "case_objects": cases,
    }


def solve_implicit_cases(specs: list[Spec], policy: str = "fp64", flux: str = "hllc", states=None, references=None,
                         max_iterations: int = 500, residual_orders: float = 8.0, start_order: str = "first",
                         workers: int = 4, chunk: int = 64, finish=None, log=print) -> tuple[list[dict], float]:
    """Each case by the steady implicit path (EulerPhysics::steady_step: pseudo-transient continuation,
    GCR and block Gauss-Seidel on the device) on a stepper and stream of its own, `workers` cases
    at a time (host threads: every run is one library call, mxeuler_implicit_steady, which waits
    only on its own stream). Per chunk of `chunk` cases the meshes are built and uploaded first,
    then solved together; the wall time counts the solve phases only (mesh set-up and upload
    excluded, as for one stepper). `finish(k, case, stepper, result)` sees each case in order, in
    the calling thread, before its stepper closes. Returns per case its result, cell count and own
    solve seconds, and the solve phases' wall seconds."""
    from concurrent.futures import ThreadPoolExecutor, wait

    from autan_amd import steady
    from autan_amd.stepper import DeviceStepper

    def setup(k):
        c = build(specs[k], policy, flux)
        dev = DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc)
        dev.upload(c.u if states is None else states[k])
        return c, dev

    def solve(k, c, dev):
        ref = None if references is None else np.asarray(references[k])
        t0 = time.perf_counter()
        r = steady.solve_implicit(dev, max_iterations=max_iterations, residual_orders=residual_orders,
                                  scales=residual_scales(c), reference=ref, start_order=start_order)
        sec = time.perf_counter() - t0
        log(f"implicit: case {k} ({c.mesh.n_cells} cells) {('converged', 'budget', 'diverged')[r.status]} in "
            f"{r.iterations} iterations, {sec:.2f} s")
        return r, sec

    out: list[dict] = []
    wall = 0.0
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        for b0 in range(0, len(specs), chunk):
            ks = range(b0, min(b0 + chunk, len(specs)))
            made = [pool.submit(setup, k) for k in ks]
            built = []
            runs = []
            try:
                for f in made:
                    built.append(f.result())
                t0 = time.perf_counter()
                runs = [pool.submit(solve, k, c, dev) for k, (c, dev) in zip(ks, built)]
                done = [f.result() for f in runs]
                sec = time.perf_counter() - t0
                wall += sec
                for k, (c, dev), (r, own) in zip(ks, built, done):
                    if finish is not None:
                        finish(k, c, dev, r)
                    out.append({"result": r, "cells": c.mesh.n_cells, "seconds": own})
            finally:
                # no stepper closes while a solve of the chunk may still use it
                wait(made)
                wait(runs)
                for f in made:
                    if f.done() and f.exception() is None:
                        f.result()[1].close()
            log(f"implicit: cases {b0}..{ks[-1]} solved in {sec:.1f} s ({workers} at a time; {wall:.1f} s so far)")
    return out, wall


def run_implicit(n_cases: int = 64, seed: int = 1, max_iterations: int = 500, residual_orders: float = 8.0,
                 policy: str = "fp32+acc64", flux: str = "hllc", cells=(2e4, 5e5), workers: int = 4,
                 log=print) -> dict:
    """The sweep by the steady implicit path, every case from the freestream to a 1e-8 relative
    residual or the fp32+acc64 round-off plateau (solve_implicit_cases, `workers` cases at a time).
    Wall time counts the solves; the meshes' set-up and their upload are excluded."""
    sp = specs(n_cases, seed, cells)
    status = ["converged", "budget", "diverged"]
    solved, wall = solve_implicit_cases(sp, policy, flux, max_iterations=max_iterations,
                                        residual_orders=residual_orders, workers=workers, chunk=len(sp), log=log)
    rows = []
    cells_iter = 0
    for k, (s, o) in enumerate(zip(sp, solved)):
        r = o["result"]
        cells_iter += o["cells"] * r.iterations
        drop = r.reference / r.residuals
        rows.append({
            "cells": o["cells"], "h": s.h, "mach": s.mach, "alpha": s.alpha, "iterations": r.iterations,
            "status": status[r.status], "first_order_steps": r.first_order_steps,
            "first_order_end": r.first_order_end, "plateau_mask": r.plateau_mask,
            "diverged_cell": r.diverged_cell, "seconds": round(o["seconds"], 3),
            "orders": [round(math.log10(v), 2) if v > 0 and math.isfinite(v) else None for v in drop],
            "max_cfl": float(r.cfl.max()) if r.iterations else None,
            "rejected_steps": int(np.sum(r.alpha == 0.0)),
        })
        log(f"sweep (implicit): case {k} {o['cells']} cells M {s.mach:.3f} h {s.h:.4f} alpha {s.alpha:+.2f}: "
            f"{status[r.status]} in {r.iterations} iterations ({r.first_order_steps} first order), "
            f"{o['seconds']:.2f} s")
    return {"cases": rows, "wall_s": wall, "cells_iter_per_s": cells_iter / wall if wall > 0 else None,
            "total_cells": sum(r["cells"] for r in rows), "workers": workers}
