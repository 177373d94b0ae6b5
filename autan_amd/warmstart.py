"""The learned warm start of the bump sweep: a network maps a cell (centroid, distance to the bump
wall) and its case (bump height, Mach number, flow angle) to an initial state (rho, u, v, w, p). It
is trained on converged states that the batch stepper computes for cases of the sweep's
generator, and evaluated on held-out cases solved to the same tolerance from the freestream and from
the network's state: every run still converges under the plain solver's rules, so the network
only changes where the iterations start.
"""

from __future__ import annotations

import hashlib
import math
import time

import numpy as np

from autan_amd import batch, initial_state, sweep

GAMMA = 1.4


def wall_distance(points: np.ndarray, h: float, w: float = 0.2, x0: float = 1.5, length: float = 3.0,
                  samples: int = 65, chunk: int = 65536) -> np.ndarray:
    """Distance from each point to the bump wall y = h exp(-((x - x0) / w)^2), 0 <= x <= length
    (a curve extruded along z). The foot x* of a point p lies within r = |f(p_x) - p_y| of p_x (the
    vertical drop is a distance), so `samples` points of [p_x - r, p_x + r] locate its basin and
    Newton steps on the stationarity condition refine it."""
    p = np.asarray(points, dtype=np.float64)[:, :2]
    out = np.empty(p.shape[0])

    def f(x):
        return h * np.exp(-(((x - x0) / w) ** 2))

    s = np.linspace(-1.0, 1.0, samples)
    for a in range(0, p.shape[0], chunk):
        px, py = p[a : a + chunk, 0], p[a : a + chunk, 1]
        r = np.abs(f(px) - py)
        lo = np.clip(px - r, 0.0, length)
        hi = np.clip(px + r, 0.0, length)
        xs = np.clip(px[:, None] + r[:, None] * s[None, :], lo[:, None], hi[:, None])
        d2 = (xs - px[:, None]) ** 2 + (f(xs) - py[:, None]) ** 2
        x = xs[np.arange(xs.shape[0]), np.argmin(d2, axis=1)]
        for _ in range(4):
            t = (x - x0) / w
            fx = h * np.exp(-t * t)
            f1 = -2.0 * t / w * fx
            f2 = (4.0 * t * t - 2.0) / (w * w) * fx
            # minimise (x - px)^2 + (f(x) - py)^2
            g = (x - px) + (fx - py) * f1
            gg = 1.0 + f1 * f1 + (fx - py) * f2
            step = g / np.where(np.abs(gg) > 1e-12, gg, 1.0)
            x_new = np.clip(x - step, lo, hi)
            # keep a step only where it does not move uphill
            better = (x_new - px) ** 2 + (f(x_new) - py) ** 2 <= (x - px) ** 2 + (f(x) - py) ** 2
            x = np.where(better, x_new, x)
        out[a : a + chunk] = np.hypot(px - x, py - f(x))
    return out


def features(mesh, spec: sweep.Spec, cells: np.ndarray | None = None) -> np.ndarray:
    """n x 7 (for every cell, or the listed ones): centroid x, y, z, distance to the bump wall, bump
    height, Mach number, flow angle."""
    x = mesh.centroids()
    if cells is not None:
        x = x[cells]
    d = wall_distance(x, spec.h)
    n = x.shape[0]
    return np.column_stack([x, d, np.full(n, spec.h), np.full(n, spec.mach), np.full(n, spec.alpha)]).astype(np.float32)


def primitive(u: np.ndarray, gamma: float = GAMMA) -> np.ndarray:
    """n x 5 (rho, u, v, w, p) of a 5 x n conserved state."""
    q = np.asarray(u).reshape(5, -1)
    rho = q[0]
    v = q[1:4] / rho
    p = (gamma - 1.0) * (q[4] - 0.5 * rho * np.sum(v * v, axis=0))
    return np.column_stack([rho, v[0], v[1], v[2], p])


def conserved(w: np.ndarray, gamma: float = GAMMA) -> np.ndarray:
    """5 x n conserved state (flat) of n x 5 primitives, as the reference's runtime builds a start
    state from an initial-state file."""
    return initial_state.conserved(w, gamma).reshape(-1)


def solve(specs: list[sweep.Spec], states=None, references=None, max_iterations=20000, residual_orders=8.0,
          policy="fp32+acc64", flux="hllc", batch_cases=64, sample=None, start_order="first", log=print):
    """Solves the cases (from the freestream, or from `states`) in batches of `batch_cases`;
    returns per case its result, cell count, wall loads and wall area, and the total seconds of the
    solves.
    `sample(index, mesh, spec, state, result)` sees each case's final state before its mesh is
    dropped."""
    out = []
    total = 0.0
    for b0 in range(0, len(specs), batch_cases):
        part = specs[b0 : b0 + batch_cases]
        cases = [sweep.build(s, policy, flux) for s in part]
        if states is not None:
            for c, u in zip(cases, states[b0 : b0 + batch_cases]):
                c.u = u
        bs = batch.BatchStepper(cases, policy)
        bs.tune("graph", 1)
        bs.upload()
        scales = np.array([sweep.residual_scales(c) for c in cases])
        refs = None if references is None else np.asarray(references[b0 : b0 + batch_cases])
        t0 = time.perf_counter()
        runs = bs.steady(max_iterations=max_iterations, residual_orders=residual_orders, scales=scales, references=refs,
                         start_order=start_order, history=False)
        sec = time.perf_counter() - t0
        total += sec
        finals = bs.download()
        wall = []
        areas = []
        for k, c in enumerate(cases):
            faces = np.nonzero(c.mesh.face_zone == 2)[0] + c.mesh.n_interior
            wall.append(bs.loads(k, faces, center=(1.5, 0.0, 0.5), p_inf=1.0 / GAMMA))
            areas.append(float(np.sum(c.mesh.face_area[faces])))
        bs.close()
        log(f"warm start: solved cases {b0}..{b0 + len(part) - 1} in {sec:.1f} s")
        for k, (r, u, c, ld, a) in enumerate(zip(runs, finals, cases, wall, areas)):
            if sample is not None:
                sample(b0 + k, c.mesh, part[k], u, r)
            out.append({"result": r, "cells": c.mesh.n_cells, "loads": ld, "wall_area": a})
        del cases, finals
    return out, total


def solve_implicit(specs: list[sweep.Spec], states=None, references=None, max_iterations=500, residual_orders=8.0,
                   policy="fp64", flux="hllc", sample=None, start_order="first", workers=4, log=print, **_):
    """As solve, by the steady implicit path (EulerPhysics::steady_step on the device), one stepper
    per case, `workers` cases at a time (sweep.solve_implicit_cases). Wall time counts the solves
    (mesh set-up and upload excluded)."""
    extra: dict[int, tuple] = {}

    def finish(k, c, dev, r):
        u = dev.download()
        faces = np.nonzero(c.mesh.face_zone == 2)[0] + c.mesh.n_interior
        loads = dev.loads(faces, center=(1.5, 0.0, 0.5), p_inf=1.0 / GAMMA)
        extra[k] = (loads, float(np.sum(c.mesh.face_area[faces])))
        if sample is not None:
            sample(k, c.mesh, specs[k], u, r)

    solved, total = sweep.solve_implicit_cases(specs, policy, flux, states=states, references=references,
                                               max_iterations=max_iterations, residual_orders=residual_orders,
                                               start_order=start_order, workers=workers, finish=finish, log=log)
    out = [{**o, "loads": extra[k][0], "wall_area": extra[k][1]} for k, o in enumerate(solved)]
    return out, total


class TrainingSet:
    """A random sample of each solved case's cells (solve's `sample` callback): features and final
    primitives."""

    def __init__(self, per_case: int = 4096, seed: int = 0) -> None:
        self.rng = np.random.default_rng(seed)
        self.per_case = per_case
        self.xs: list[np.ndarray] = []
        self.ts: list[np.ndarray] = []
        self.skipped = 0
        self.converged = 0
        self.left_out: list[dict] = []

    def __call__(self, index, mesh, spec, state, result) -> None:
        # A case that diverged teaches nothing: it is left out and reported (`left_out`).
        if result.status == 2 or not np.all(np.isfinite(state)):
            self.skipped += 1
            self.left_out.append({"case": int(index), "h": spec.h, "mach": spec.mach, "alpha": spec.alpha,
                                  "cells": int(mesh.n_cells), "iterations": int(result.iterations),
                                  "diverged_cell": result.diverged_cell})
            return
        self.converged += int(result.status == 0)
        n = mesh.n_cells
        pick = np.sort(self.rng.choice(n, size=min(self.per_case, n), replace=False))
        self.xs.append(features(mesh, spec, pick))
        self.ts.append(primitive(state)[pick].astype(np.float32))

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        return np.concatenate(self.xs), np.concatenate(self.ts)


def warm_states(net, specs: list[sweep.Spec], policy="fp32+acc64", flux="hllc"):
    """The network's initial state for each case (density and pressure kept positive), and the
    seconds spent in the network's inputs and inference (the meshes' set-up excluded: a freestream
    start needs it too)."""
    out = []
    seconds = 0.0
    for s in specs:
        c = sweep.build(s, policy, flux)
        t0 = time.perf_counter()
        out.append(conserved(predicted(net, c.mesh, s)))
        seconds += time.perf_counter() - t0
    return out, seconds


def predicted(net, mesh, spec: sweep.Spec) -> np.ndarray:
    """n x 5 primitives (rho, u, v, w, p) the network gives the mesh's cells, density and pressure
    kept positive."""
    w = net.predict(features(mesh, spec)).astype(np.float64)
    w[:, 0] = np.maximum(w[:, 0], 0.05)
    w[:, 4] = np.maximum(w[:, 4], 0.05 / GAMMA)
    return w


def mesh_digest(mesh) -> str:
    """SHA-256 of a generated mesh's points and faces (the reference hashes a mesh's file; a mesh
    built in memory has none)."""
    h = hashlib.sha256()
    for name in ("x", "y", "z", "face_node_offsets", "face_nodes", "owner", "neighbour"):
        h.update(np.ascontiguousarray(mesh.arrays[name]).tobytes())
    return h.hexdigest()


def write_initial_state(path, net, spec: sweep.Spec, policy="fp32+acc64", flux="hllc") -> tuple[str, object]:
    """The network's state of a sweep case as an autan.initial-state/1 file (one row per cell at its
    centroid), the warm start's interchange with the reference's SDK and runtime. Returns the
    file's SHA-256 and the case."""
    c = sweep.build(spec, policy, flux)
    digest = initial_state.write(path, oracle=f"autan_amd warm start: h={spec.h!r} M={spec.mach!r} alpha={spec.alpha!r}",
                                 mesh_sha256=mesh_digest(c.mesh), centroids=c.mesh.centroids(),
                                 values=predicted(net, c.mesh, spec))
    return digest, c


def train_network(specs: list[sweep.Spec], iterations: int, epochs: int, per_case: int = 4096, batch: int = 4096,
                  lr: tuple[float, float] = (1e-3, 1e-5), seed: int = 0, log=print, method: str = "explicit",
                  policy: str = "fp32+acc64", workers: int = 4):
    """The network trained on the training cases' states: each case solved from the freestream
    (explicit: `iterations` steady iterations of the batch stepper; implicit: the steady implicit
    path to a 1e-8 relative residual or its round-off plateau, at most `iterations`), cases that
    diverge left out and listed, `per_case` random cells of each kept, then `epochs` of Adam with a
    cosine learning rate from lr[0] to lr[1]. Returns the network and what it took."""
    from autan_amd.mlp import MLP

    collect = TrainingSet(per_case=per_case, seed=seed)
    solver = solve_implicit if method == "implicit" else solve
    kw = {"workers": workers} if method == "implicit" else {}
    _, t_gen = solver(specs, max_iterations=iterations, residual_orders=8.0, sample=collect, log=log, policy=policy, **kw)
    x, t = collect.arrays()
    log(f"training set: {x.shape[0]} cells of {len(specs) - collect.skipped} cases ({collect.skipped} diverged, "
        f"left out; {t_gen:.1f} s of solves)")
    net = MLP(7, 256, 4, 5, seed=seed)
    net.fit_normalization(x, t)
    t_train = 0.0
    first = loss = float("nan")
    for ep in range(epochs):
        rate = lr[1] + (lr[0] - lr[1]) * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
        loss, sec = net.train(x, t, epochs=1, batch=batch, lr=rate, seed=seed + ep)
        if ep == 0:
            first = loss
        t_train += sec
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"epoch {ep}: loss {loss:.3e} ({sec:.2f} s)")
    return net, {"cells": int(x.shape[0]), "diverged_left_out": collect.skipped, "left_out": collect.left_out,
                 "converged": collect.converged, "method": method, "policy": policy, "solve_s": round(t_gen, 1),
                 "train_s": round(t_train, 2), "first_loss": first, "loss": loss}


def compare(net, specs: list[sweep.Spec], orders: float, max_iterations: int, log=print, method: str = "explicit",
            policy: str = "fp32+acc64", workers: int = 4) -> dict:
    """The held-out cases solved from the freestream (first-order start) and from the network's state
    (second order from the first step), the latter both against its own first residual and against
    the freestream run's (the absolute threshold a warm start from another run is held to)."""
    solver = solve_implicit if method == "implicit" else solve
    kw = {"workers": workers} if method == "implicit" else {}
    free, t_free = solver(specs, max_iterations=max_iterations, residual_orders=orders, log=log, policy=policy, **kw)
    warm_u, t_infer = warm_states(net, specs, policy=policy)
    warm, t_warm = solver(specs, states=warm_u, max_iterations=max_iterations, residual_orders=orders,
                          start_order="second", log=log, policy=policy, **kw)
    refs = [f["result"].reference for f in free]
    warm_abs, t_warm_abs = solver(specs, states=warm_u, references=refs, max_iterations=max_iterations,
                                  residual_orders=orders, start_order="second", log=log, policy=policy, **kw)
    return {"free": free, "warm": warm, "warm_abs": warm_abs, "free_s": t_free, "warm_s": t_warm,
            "warm_abs_s": t_warm_abs, "inference_s": t_infer}
