"""Benchmark: SSP-RK3 steps of the explicit Euler residual on one MI355X.

Default workload (the roofline target: second-order HLLC in fp32+acc64 on about 8 M polyhedral
cells): the hybrid bump channel, a 3 x 1 x 1 channel with a Gaussian bump (height 0.1, width 0.2)
on its lower wall, meshed as a 462 x 154 x 154 hexahedral lattice with prism rows on the wall and a
band of merged polyhedra with non-planar faces (7,915,215 cells), cells renumbered along a Morton
curve; inflow Mach 0.5 from the freestream start, slip walls, characteristic farfield; face2
least-squares gradients with the Venkatakrishnan limiter, SSP-RK3 on local time steps (CFL 0.8).
One step = one SSP-RK3 step (three residuals, the norms of R(u), the spectral radius and time
steps, validity checks). The run also reports the residual drop per second of a steady run from the
freestream by the explicit steps and by the steady implicit path, the conservation check and the
fp32+acc64 residual's distance from the fp64 one on the same state.

Other workloads (--workload): `tets`, the manufactured-solution Kuhn tetrahedra (128^3 x 6 =
12,582,912 cells, jittered, HLLC, global CFL 0.5); `vortex`, the isentropic vortex on 32^3 periodic
hexahedra (first-order Roe, fp64, 200 steps, compared with the CPU oracle's 200 steps); `sweep`,
`warmstart`, `mms`, `mlp` (their own lines).

Prints one JSON line: ns per cell per step (lower is better), the roofline of the dominant kernel
(algorithmic bytes per launch over its event-timed duration, against the 8 TB/s HBM peak), and a
CPU baseline: the oracle (this repo's C++ restatement of the reference's CPU loops) timed on the
same mesh and state for a bounded number of steps on this host, serial and on every host thread the
process may use.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    """The host threads this process may use: OMP_NUM_THREADS where set, else its CPU affinity,
    at most 16 (one GPU's share of the host)."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_workload(name: str, args, size=None):
    from autan_amd import workloads

    t0 = time.time()
    if name == "tets":
        w = workloads.tets(size or args.n, args.policy, args.flux, args.order)
    elif name == "vortex":
        w = workloads.vortex(size or 32, args.policy, args.flux)
    else:
        nx, ny, nz = size or (462, 154, 154)
        w = workloads.bump(nx, ny, nz, args.policy, args.flux, order=args.order)
    m = w.mesh
    log(f"{name}: {m.n_cells} cells, {m.n_faces} faces, {m.n_stencil} stencil entries ({time.time() - t0:.1f} s)")
    return w


def cpu_baseline(w, steps: int, sample: str) -> tuple[dict, object]:
    """The oracle (the reference CPU path's semantics: serial gather loops) on a bounded sample,
    then the same loops on every host thread this process may use. Returns the record and the
    state after the serial steps."""
    import numpy as np

    from oracle import oracle as orc
    from oracle.oracle import Oracle

    m = w.mesh
    ora = Oracle(m, w.scheme, w.policy, w.bc_kind, w.bc, w.source)
    timings, states = {}, {}
    for threads in dict.fromkeys((1, host_threads())):
        orc.set_threads(threads)
        u = w.u.copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            u, norms, h, inv = ora.step(u, cfl=w.cfl, local=w.local)
        timings[threads] = time.perf_counter() - t0
        states[threads] = u
        if inv is not None or not np.all(np.isfinite(norms)):
            raise RuntimeError(f"the oracle diverged: invalid cell {inv}")
    orc.set_threads(1)
    t1 = timings[1]
    tn = max(timings)
    record = {
        "value": round(t1 / steps / m.n_cells * 1e9, 2),
        "unit": "ns/cell/iter",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} SSP-RK3 steps of {sample} ({m.n_cells} cells), serial oracle (gather loops), {t1:.1f} s",
        "threaded": {
            "value": round(timings[tn] / steps / m.n_cells * 1e9, 2),
            "cores": tn,
            "cpu": cpu_model(),
            "seconds": round(timings[tn], 2),
        },
    }
    return record, states[1]


def stream_peak() -> dict:
    from autan_amd import _lib

    lib = _lib.euler()
    gbps = (ctypes.c_double * 4)()
    ms = (ctypes.c_double * 4)()
    err = ctypes.create_string_buffer(1024)
    rc = lib.mxeuler_stream(ctypes.c_int64(1 << 28), ctypes.c_int32(10), gbps, ms, err, ctypes.c_size_t(1024))
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return {"copy": gbps[0], "scale": gbps[1], "add": gbps[2], "triad": gbps[3]}


def timed_steps(dev, w, warmup: int, steps: int):
    import numpy as np

    for _ in range(warmup):
        r = dev.step(cfl=w.cfl, local=w.local)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = dev.step(cfl=w.cfl, local=w.local)
    elapsed = time.perf_counter() - t0
    if r.invalid is not None or not np.all(np.isfinite(r.residuals)):
        raise RuntimeError(f"the run diverged: invalid cell {r.invalid}, norms {r.residuals}")
    return elapsed, r


def roofline(dev, w, ms_step: float, tag: str) -> tuple[dict, dict]:
    from autan_amd import traffic

    prof = dev.profile(steps=5, cfl=w.cfl, local=w.local)
    counts = traffic.Counts.of(w.mesh, **w.counts)
    kb = traffic.kernel_bytes(counts, w.policy, w.local)
    per_step = {k: v["ms"] * v["launches"] for k, v in prof["kernels"].items()}
    dominant = max(per_step, key=per_step.get)
    ms_launch = prof["kernels"][dominant]["ms"]
    achieved = kb[dominant] / (ms_launch * 1e-3) / 1e9 if dominant in kb else None
    kernels = {
        k: {
            "ms": round(v["ms"], 4),
            "launches": v["launches"],
            "GB/s": round(kb[k] / (v["ms"] * 1e-3) / 1e9, 1) if k in kb and v["ms"] > 0 else None,
        }
        for k, v in prof["kernels"].items()
    }
    step_b = traffic.step_bytes(counts, w.policy, w.local)
    ref_b = traffic.reference_step_bytes(counts, w.policy, w.local)
    peak = stream_peak()
    log(f"profile: {json.dumps(kernels)}")
    log(f"STREAM: {json.dumps(peak)}")
    traffic_file = ROOT / "profiles" / f"traffic_{dominant}_{tag}_{w.policy}.json"
    measured = json.loads(traffic_file.read_text())["bytes_per_launch"] if traffic_file.exists() else None
    roof = {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": None if achieved is None else round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured,
        "algorithmic_bytes_per_launch": kb.get(dominant),
        "ms_per_launch": round(ms_launch, 4),
    }
    step = {
        "event_ms": round(prof["ms_per_step"], 4),
        "algorithmic_bytes": step_b,
        "achieved_GBs": round(step_b / (ms_step * 1e-3) / 1e9, 1),
        "frac_of_peak": round(step_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "reference_B_step": ref_b,
        "reference_B_step_frac": round(ref_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stream_GBs": {k: round(v, 1) for k, v in peak.items()},
        "kernels": kernels,
    }
    return roof, step


def flux_scale(w):
    """f_k A_i (5, n): the reference's per-cell error unit, f = (rho q, rho q^2 x 3, rho H q) of the
    freestream with q = |v| + c, A_i the cell's surface."""
    import numpy as np

    ref = np.array(w.scheme.reference)
    g = w.scheme.gamma
    v = ref[1:4] / ref[0]
    p = (g - 1.0) * (ref[4] - 0.5 * ref[0] * (v @ v))
    c = math.sqrt(g * p / ref[0])
    q = math.sqrt(v @ v) + c
    h = c * c / (g - 1.0) + 0.5 * (v @ v)
    f = np.array([ref[0] * q, ref[0] * q * q, ref[0] * q * q, ref[0] * q * q, ref[0] * h * q])
    m = w.mesh
    area = np.zeros(m.n_cells)
    np.add.at(area, m.owner, m.face_area)
    np.add.at(area, m.neighbour, m.face_area[: m.n_interior])
    return f[:, None] * area[None, :]


def bump_extras(dev, w, args) -> dict:
    """The bump channel's steady run (orders of residual drop per second from the freestream
    start), its conservation check and the fp32+acc64 residual against the fp64 one."""
    import numpy as np

    from autan_amd import steady
    from autan_amd.stepper import DeviceStepper

    out = {}
    dev.upload(w.u)
    ref = np.array(w.scheme.reference)
    speed = float(np.linalg.norm(ref[1:4])) / ref[0]
    p = (w.scheme.gamma - 1.0) * (ref[4] - 0.5 * ref[0] * speed * speed)
    scales = dev.residual_scales(ref[0], speed, p)
    t0 = time.perf_counter()
    run = steady.solve(dev, max_iterations=args.steady_iterations, residual_orders=8.0, scales=scales)
    t = time.perf_counter() - t0
    drop = run.reference / run.residuals
    out["steady"] = {
        "iterations": run.iterations,
        "seconds": round(t, 3),
        "first_order_steps": run.first_order_steps,
        "first_order_end": run.first_order_end,
        "orders_dropped": [round(math.log10(d), 3) if d > 0 and math.isfinite(d) else None for d in drop],
        "mass_orders_per_second": round(math.log10(drop[0]) / t, 3) if drop[0] > 0 else None,
        "status": ["converged", "budget", "diverged"][run.status],
    }
    c = dev.conservation()
    out["conservation"] = {
        "mass_E": c["E"][0],
        "mass_E_over_D": c["E"][0] / c["D"][0],
        "E": [float(v) for v in c["E"]],
        "D": [float(v) for v in c["D"]],
    }
    if w.policy == "fp32+acc64" and not args.no_fp64_check:
        u = dev.download()
        r32 = dev.residual(u)
        dev64 = DeviceStepper(w.mesh, w.scheme, "fp64", w.bc_kind, w.bc, w.source)
        r64 = dev64.residual(u)
        dev64.close()
        err = np.abs(r32 - r64).reshape(5, -1) / flux_scale(w)
        out["fp32_vs_fp64"] = {
            "state": f"after the {run.iterations}-iteration steady run",
            "max_err_over_f_A": [float(e) for e in err.max(axis=1)],
            "max_in_eps32": round(float(err.max() / np.finfo(np.float32).eps), 2),
        }
    if args.implicit_iterations > 0:
        # The same channel from the same freestream state by the steady implicit path
        # (pseudo-transient continuation, GCR + multicolour block Gauss-Seidel on the device).
        dev.upload(w.u)
        t0 = time.perf_counter()
        imp = steady.solve_implicit(dev, max_iterations=args.implicit_iterations, residual_orders=8.0,
                                    scales=scales)
        t = time.perf_counter() - t0
        drop = imp.reference / imp.residuals
        out["steady_implicit"] = {
            "iterations": imp.iterations,
            "seconds": round(t, 3),
            "first_order_steps": imp.first_order_steps,
            "orders_dropped": [round(math.log10(d), 3) if d > 0 and math.isfinite(d) else None for d in drop],
            "mass_orders_per_second": round(math.log10(drop[0]) / t, 3) if drop[0] > 0 else None,
            "gcr_directions": int(np.sum(imp.cycles)),
            "max_cfl": float(imp.cfl.max()) if imp.iterations else None,
            "status": ["converged", "budget", "diverged"][imp.status],
        }
        log(f"implicit steady run: {imp.iterations} iterations in {t:.1f} s, orders {out['steady_implicit']['orders_dropped']}")
    dev.close()
    return out


def sweep_implicit_main(args) -> int:
    """The sweep by the steady implicit path: every case from the freestream to a 1e-8 relative
    residual (fp64) or the round-off plateau (fp32+acc64), one stepper per case, --workers at a time; per-case
    iterations and seconds, and the whole sweep's wall time."""
    import numpy as np

    from autan_amd import sweep

    policy = args.policy or "fp32+acc64"
    t0 = time.time()
    rep = sweep.run_implicit(n_cases=args.cases, seed=1, max_iterations=args.max_iterations, policy=policy,
                             flux=args.flux or "hllc", workers=args.workers, log=log)
    cases = rep["cases"]
    iters = [c["iterations"] for c in cases]
    line = {
        "metric": "wall time to 1e-8 relative residual (or the round-off plateau), whole sweep",
        "value": round(rep["wall_s"], 2),
        "unit": "s",
        "n_gpus": 1,
        "steps": int(max(iters)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc" if policy == "fp32+acc64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.cases} hybrid bump channels (seed 1): cells log-uniform in [2e4, 5e5], h U(0.02, 0.12), "
            f"Mach U(0.3, 0.8), alpha U(-4, 4) deg; second-order HLLC {policy}, steady implicit path (pseudo-transient "
            f"continuation, GCR + symmetric multicolour block Gauss-Seidel on the device), first-order start, "
            f"to a 1e-8 drop or the round-off plateau",
            "cells": rep["total_cells"],
        },
        "cells_iter_per_s": round(rep["cells_iter_per_s"], 1),
        "workers": rep["workers"],
        "case_seconds_sum": round(sum(c["seconds"] for c in cases), 2),
        "converged": sum(c["status"] == "converged" for c in cases),
        "diverged": sum(c["status"] == "diverged" for c in cases),
        "iterations": {"min": int(min(iters)), "median": float(np.median(iters)), "max": int(max(iters))},
        "cases": cases,
        "setup_s": round(time.time() - t0 - rep["wall_s"], 1),
    }
    print(json.dumps(line), flush=True)
    return 0


def sweep_main(args) -> int:
    """The batched sweep: every case stepped together to its steady state; aggregate cells x
    iterations per second over the whole run (one batch, HIP graph replay, autotuned variant)."""
    import numpy as np

    from autan_amd import sweep

    if args.method == "implicit":
        return sweep_implicit_main(args)

    t0 = time.time()
    rep = sweep.run(n_cases=args.cases, seed=1, max_iterations=args.max_iterations, policy=args.policy or "fp32+acc64",
                    flux=args.flux or "hllc", log=log)
    cases = rep["cases"]
    iters = [c["iterations"] for c in cases]
    line = {
        "metric": "ns/cell/iter",
        "value": round(1e9 / rep["cells_iter_per_s"], 4),
        "unit": "ns/cell/iter",
        "n_gpus": 1,
        "steps": int(max(iters)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc",
        "data": "synthetic",
        "config": {
            "workload": f"{args.cases} hybrid bump channels (seed 1): cells log-uniform in [2e4, 5e5], h U(0.02, 0.12), "
            f"Mach U(0.3, 0.8), alpha U(-4, 4) deg; second-order HLLC fp32+acc64, first-order start, local time steps "
            f"CFL 0.8, to a 1e-8 drop or the round-off plateau; one batch stepper, HIP graph replay",
            "cells": rep["total_cells"],
        },
        "cells_iter_per_s": round(rep["cells_iter_per_s"], 1),
        "wall_s": round(rep["wall_s"], 2),
        "tuned": rep["tuned"],
        "converged": sum(c["status"] == "converged" for c in cases),
        "iterations": {"min": int(min(iters)), "median": float(np.median(iters)), "max": int(max(iters))},
        "cases": cases,
        "setup_s": round(time.time() - t0 - rep["wall_s"], 1),
    }
    print(json.dumps(line), flush=True)
    return 0


def warmstart_main(args) -> int:
    """The learned warm start: run the training cases (the sweep's generator, seed 11, meshes of
    2e4 to 1e5 cells) for a fixed number of iterations, train the network on their cells' states,
    then solve the held-out cases (seed 12, the sweep's full range) from the freestream and from the
    network's state to the same absolute residual (the freestream run's iteration-1 residual times
    10^-orders, the threshold a warm start from another run is held to)."""
    import numpy as np

    from autan_amd import sweep, warmstart

    t0 = time.time()
    train_specs = sweep.specs(args.train_cases, seed=11, cells=(2e4, 1e5))
    test_specs = sweep.specs(args.test_cases, seed=12)
    implicit = args.method == "implicit"
    policy = args.policy or ("fp64" if implicit else "fp32+acc64")
    # The implicit path runs to the configuration's 1e-8 within a few hundred iterations; the
    # explicit one stops at 2 orders (its 20000-iteration budget reaches no further).
    if args.test_orders is None:
        args.test_orders = 8.0 if implicit else 2.0
    net, train = warmstart.train_network(train_specs, args.train_iterations if not implicit else args.max_iterations,
                                         args.epochs, log=log, method=args.method, policy=policy,
                                         workers=args.workers)
    fwd_ms = net.time_forward(1 << 18, 5)
    tflops = net.flops_per_sample() * (1 << 18) / (fwd_ms * 1e-3) / 1e12
    log(f"network: loss {train['loss']:.3e} after {args.epochs} epochs ({train['train_s']:.1f} s); forward "
        f"{tflops:.1f} TFLOP/s")
    cmp = warmstart.compare(net, test_specs, args.test_orders, args.max_iterations, log=log, method=args.method,
                            policy=policy, workers=args.workers)
    free, warm, warm_abs = cmp["free"], cmp["warm"], cmp["warm_abs"]
    t_free, t_warm, t_warm_abs, t_infer = cmp["free_s"], cmp["warm_s"], cmp["warm_abs_s"], cmp["inference_s"]
    rows = []
    for s, f, w, wa in zip(test_specs, free, warm, warm_abs):
        ff = f["loads"][0]
        dw = float(np.linalg.norm(wa["loads"][0] - ff) / max(np.linalg.norm(ff), 1e-300))
        # force coefficients: the wall force over the freestream dynamic pressure (rho = 1,
        # p = 1/gamma: q = M^2 / 2) times the wall's area
        q_area = 0.5 * s.mach * s.mach * f["wall_area"]
        dc = float(np.linalg.norm(wa["loads"][0] - ff) / q_area)
        rows.append({
            "cells": f["cells"], "h": s.h, "mach": s.mach, "alpha": s.alpha,
            "freestream": [f["result"].iterations, ["converged", "budget", "diverged"][f["result"].status]],
            "warm_own_reference": [w["result"].iterations, ["converged", "budget", "diverged"][w["result"].status]],
            "warm_freestream_reference": [wa["result"].iterations,
                                          ["converged", "budget", "diverged"][wa["result"].status]],
            "force_free": [float(v) for v in ff], "force_rel_diff": dw, "force_coeff_diff": dc,
            "seconds": [f.get("seconds"), wa.get("seconds")],
        })
    it_free = sum(r["freestream"][0] for r in rows)
    it_warm = sum(r["warm_freestream_reference"][0] for r in rows)
    # The cases both starts bring to the threshold: their iterations, own seconds and loads.
    both = [r for r in rows if r["freestream"][1] == "converged" and r["warm_freestream_reference"][1] == "converged"]
    converged = {
        "cases": len(both),
        "iterations_freestream": sum(r["freestream"][0] for r in both),
        "iterations_warm": sum(r["warm_freestream_reference"][0] for r in both),
        "max_force_coeff_diff": max((r["force_coeff_diff"] for r in both), default=None),
    }
    if both and all(r["seconds"][0] is not None for r in both):
        converged["seconds_freestream"] = round(sum(r["seconds"][0] for r in both), 2)
        converged["seconds_warm"] = round(sum(r["seconds"][1] for r in both), 2)
    line = {
        "metric": "wall time to the residual threshold, learned vs freestream",
        "value": round(t_free / (t_warm_abs + t_infer), 3),
        "unit": "x (freestream seconds / warm seconds)",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": f"bf16 (network), {'f32+f64acc' if policy == 'fp32+acc64' else 'f64'} (solver)",
        "data": "synthetic",
        "config": {
            "workload": f"MLP 7-256x4-5 SiLU, BF16 matrix cores, trained on {len(train_specs)} bump channels "
            f"(seed 11, 2e4-1e5 cells) "
            + (f"solved by the steady implicit path ({policy}) to a 1e-8 drop or the round-off plateau"
               if implicit else f"after {args.train_iterations} iterations each")
            + f", tested on {len(test_specs)} held-out channels (seed 12, 2e4-5e5 cells) solved "
            f"{'by the steady implicit path ' if implicit else ''}to {args.test_orders:g} orders below the "
            f"freestream run's first residual",
        },
        "train": {**train, "forward_TFLOPs": round(tflops, 1), "forward_ms_262144": round(fwd_ms, 3)},
        "test": {"freestream_s": round(t_free, 2), "warm_s_own_reference": round(t_warm, 2),
                 "warm_s_freestream_reference": round(t_warm_abs, 2), "inference_s": round(t_infer, 2),
                 "iterations_freestream": it_free, "iterations_warm": it_warm,
                 "max_force_rel_diff": max(r["force_rel_diff"] for r in rows),
                 "max_force_coeff_diff": max(r["force_coeff_diff"] for r in rows),
                 "converged_from_both": converged},
        "cases": rows,
        "total_s": round(time.time() - t0, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


def mlp_main(args) -> int:
    """The warm-start network's kernels alone: the fused BF16 inference on device-resident samples
    (against the dense BF16 matrix-core peak) and one training epoch (device-resident data, the
    recorded step replayed per batch), on synthetic inputs of the network's shape."""
    import numpy as np

    from autan_amd.mlp import MLP

    rng = np.random.default_rng(0)
    net = MLP(7, 256, 4, 5, seed=0)
    n_train = 1 << 21
    x = rng.uniform(-1, 1, size=(n_train, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1], np.cos(x[:, 2] * x[:, 3]), x[:, 4] ** 2, x[:, 5] - x[:, 6],
                  np.tanh(x[:, 0] + x[:, 6])], axis=1).astype(np.float32)
    net.fit_normalization(x, t)
    # The inference tile (samples per workgroup) with the shortest time on 2^20 samples.
    tiles = {}
    for rows in (64, 96, 128):
        net.set_infer_rows(rows)
        ms = net.time_forward(1 << 20, 20)
        tiles[str(rows)] = round(net.flops_per_sample() * (1 << 20) / (ms * 1e-3) / 1e12, 1)
        log(f"inference tile {rows} samples per workgroup: {ms:.3f} ms for 2^20 samples, {tiles[str(rows)]} TFLOP/s")
    best = int(max(tiles, key=lambda k: tiles[k]))
    net.set_infer_rows(best)
    infer = {"tiles_TFLOPs": tiles, "rows": best}
    for n in (1 << 18, 1 << 20):
        ms = net.time_forward(n, 20)
        infer[str(n)] = {"ms": round(ms, 4), "TFLOP/s": round(net.flops_per_sample() * n / (ms * 1e-3) / 1e12, 1)}
        log(f"inference of {n} samples: {ms:.3f} ms, {infer[str(n)]['TFLOP/s']} TFLOP/s")
    if args.no_train:
        sec = loss = train_tf = None
    else:
        net.train(x, t, epochs=1, batch=4096, lr=1e-3, seed=0)  # records the step
        loss, sec = net.train(x, t, epochs=1, batch=4096, lr=1e-3, seed=1)
        train_tf = 3 * net.flops_per_sample() * (n_train // 4096) * 4096 / sec / 1e12
        log(f"training epoch of {n_train} samples: {sec:.3f} s, {train_tf:.1f} TFLOP/s effective, loss {loss:.3e}")
    big = infer[str(1 << 20)]
    line = {
        "metric": "BF16 inference TFLOP/s (real products) of the warm-start network",
        "value": big["TFLOP/s"],
        "unit": "TFLOP/s",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "MLP 7-256x4-5 SiLU, fused inference of 2^20 samples; training epoch of 2^21 "
                               "samples at batch 4096"},
        "roofline": {"bound": "mfma", "kernel": "k_infer", "achieved": big["TFLOP/s"], "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(big["TFLOP/s"] / 2500.0, 4), "traffic": None},
        "inference": infer,
        "training_epoch": None if sec is None else {"seconds": round(sec, 3), "TFLOP/s_effective": round(train_tf, 1),
                                                    "loss": loss},
    }
    print(json.dumps(line), flush=True)
    return 0


def mms_main(args) -> int:
    """Observed order of accuracy on the manufactured solution (Kuhn tetrahedra at the listed
    levels; steady discrete solutions on the device, errors at the centroids)."""
    from autan_amd import order

    t0 = time.time()
    levels = tuple(int(v) for v in args.levels.split(","))
    policies = tuple(args.policies.split(","))
    rep = order.study(levels, policies, flux=args.flux or "hllc", limiter=not args.unlimited,
                      max_iterations=args.max_iterations, residual_orders=args.residual_orders, log=log)
    line = {
        "metric": "observed order of accuracy (manufactured solution, volume-weighted L2 of the conserved variables)",
        "value": None,
        "unit": "order",
        "n_gpus": 1,
        "dtype": ",".join(policies),
        "data": "synthetic",
        "config": {"workload": f"Kuhn tetrahedra {levels} ^3 x 6 (jitter 10 %, seed 0), manufactured solution, "
                               f"second-order {(args.flux or 'hllc').upper()}, face2 LSQ, "
                               f"{'no limiter' if args.unlimited else 'Venkatakrishnan K=5'}, steady runs on local time "
                               f"steps from the exact state to {args.residual_orders:g} orders or "
                               f"{args.max_iterations} iterations"},
        **rep,
        "total_s": round(time.time() - t0, 1),
    }
    fitted = [v[0] for v in rep["orders"].get("fp32+acc64", rep["orders"].get("fp64", {})).values()]
    line["value"] = round(min(fitted), 3) if fitted else None
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (tets 20, bump 20, vortex 200)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bump",
                    choices=["tets", "vortex", "bump", "sweep", "warmstart", "mms", "mlp"])
    ap.add_argument("--levels", default="64,96,128", help="mms: tetrahedra levels n (n^3 x 6 cells)")
    ap.add_argument("--no-train", action="store_true", help="mlp: inference only (the training step is a recorded "
                    "HIP graph, which kernel tracing does not follow here)")
    ap.add_argument("--policies", default="fp64,fp32+acc64", help="mms: policies")
    ap.add_argument("--unlimited", action="store_true", help="mms: no limiter (the design order)")
    ap.add_argument("--residual-orders", type=float, default=8.0, help="mms: relative residual drop of the steady runs")
    ap.add_argument("--train-cases", type=int, default=512, help="warmstart: solved cases the network learns from")
    ap.add_argument("--test-cases", type=int, default=32, help="warmstart: held-out cases")
    ap.add_argument("--train-iterations", type=int, default=8000, help="warmstart: iterations of the training solves")
    ap.add_argument("--test-orders", type=float, default=None,
                    help="warmstart: residual drop the held-out solves run to, from the freestream run's reference "
                    "(default 8 with --method implicit, 2 with the explicit method)")
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--method", default="explicit", choices=["explicit", "implicit"],
                    help="sweep / warmstart: the batch stepper's explicit SSP-RK3 steps, or the steady implicit path "
                    "(--max-iterations is then its iteration budget per case)")
    ap.add_argument("--cases", type=int, default=64, help="sweep: number of cases")
    ap.add_argument("--workers", type=int, default=4,
                    help="sweep / warmstart with --method implicit: cases solved at a time (one stepper and stream each)")
    ap.add_argument("--max-iterations", type=int, default=None,
                    help="sweep / warmstart / mms: iteration budget per case (default 20000, 600 with --method implicit)")
    ap.add_argument("--n", type=int, default=128, help="tets: cubes per side (6 n^3 tetrahedra)")
    ap.add_argument("--bump-size", default="462,154,154", help="bump: lattice nx,ny,nz")
    ap.add_argument("--policy", default=None, choices=["fp32+acc64", "fp64"])
    ap.add_argument("--flux", default=None, choices=["roe", "hllc", "ausm+up"])
    ap.add_argument("--cpu-n", type=int, default=48, help="tets: cubes per side of the CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=None, help="steps of the CPU baseline (bump 1 on the same "
                    "mesh, tets 4 on the --cpu-n sample)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--steady-iterations", type=int, default=400)
    ap.add_argument("--implicit-iterations", type=int, default=60,
                    help="bump: budget of a steady implicit run from the freestream after the explicit one (0: none)")
    ap.add_argument("--order", default="morton", choices=["morton", "rcm"], help="cell renumbering of the set-up")
    ap.add_argument("--tune", default="", help="kernel variants, key=value,... (face_stage, graph, grad_occupancy, "
                    "face_occupancy); set after the autotuner")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel variants")
    ap.add_argument("--no-extras", action="store_true", help="bump: skip the steady run, conservation and fp64 check")
    args = ap.parse_args()
    if args.max_iterations is None:
        args.max_iterations = 600 if args.method == "implicit" else 20000
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the stepper holds a whole mesh on one device and has no "
              "multi-GPU path. Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.workload == "sweep":
        return sweep_main(args)
    if args.workload == "warmstart":
        return warmstart_main(args)
    if args.workload == "mms":
        return mms_main(args)
    if args.workload == "mlp":
        return mlp_main(args)
    defaults = {
        "tets": ("fp32+acc64", "hllc", 20),
        "vortex": ("fp64", "roe", 200),
        "bump": ("fp32+acc64", "hllc", 20),
    }[args.workload]
    args.policy = args.policy or defaults[0]
    args.flux = args.flux or defaults[1]
    steps = args.steps or defaults[2]

    import numpy as np

    from autan_amd.stepper import DeviceStepper

    size = tuple(int(v) for v in args.bump_size.split(",")) if args.workload == "bump" else None
    w = make_workload(args.workload, args, size)
    m = w.mesh
    dev = DeviceStepper(m, w.scheme, w.policy, w.bc_kind, w.bc, w.source)
    dev.upload(w.u)
    tuned = None
    if not args.no_autotune:
        tuned = dev.autotune(steps=3, cfl=w.cfl, local=w.local)
        log(f"autotune: {tuned}")
    for kv in filter(None, args.tune.split(",")):
        key, value = kv.split("=")
        dev.tune(key, int(value))
    log("uploaded; warming up")
    warmup = 0 if args.workload == "vortex" else args.warmup
    elapsed, r = timed_steps(dev, w, warmup, steps)
    ms_step = elapsed / steps * 1e3
    ns_cell = ms_step * 1e6 / m.n_cells
    log(f"{steps} steps: {ms_step:.3f} ms/step, {ns_cell:.3f} ns/cell/step, norms {r.residuals}")
    u_gpu = dev.download()
    tag = {"tets": f"tets{args.n}", "vortex": "vortex32", "bump": "bump" + "x".join(map(str, size or ()))}[args.workload]
    roof, step = roofline(dev, w, ms_step, tag)

    extras = {}
    cpu = None
    if args.workload == "vortex":
        # The CPU oracle runs the same steps from the same state: the GPU's state must be its bits.
        if not args.no_cpu:
            log("CPU oracle: the same steps")
            cpu, u_cpu = cpu_baseline(w, steps, w.description)
            d = np.abs(u_gpu - u_cpu)
            extras["parity"] = {"steps": steps, "max_abs_diff": float(d.max()), "bitwise": bool(np.array_equal(u_gpu, u_cpu))}
    elif args.workload == "bump":
        extras = {} if args.no_extras else bump_extras(dev, w, args)
        if not args.no_cpu:
            log("CPU baseline (oracle) on the same mesh and state")
            cpu, _ = cpu_baseline(w, args.cpu_steps or 1, "the same mesh from the same freestream state")
    elif not args.no_cpu:
        log("CPU baseline (oracle)")
        small = make_workload("tets", args, args.cpu_n)
        cpu, _ = cpu_baseline(small, args.cpu_steps or 4, f"the same workload at {args.cpu_n}^3 x 6 tets")

    line = {
        "metric": "ns/cell/iter",
        "value": round(ns_cell, 4),
        "unit": "ns/cell/iter",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc" if w.policy == "fp32+acc64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": w.description,
            "cells": m.n_cells,
            "faces": m.n_faces,
            "stencil_entries": m.n_stencil,
        },
        "cells_iter_per_s": round(m.n_cells / (ms_step * 1e-3), 1),
        "variant": None if tuned is None else tuned.split(";")[0].removeprefix("best "),
        "roofline": roof,
        "step": step,
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        line["gpu_vs_cpu"] = {"serial": round(cpu["value"] / ns_cell, 1),
                              "threaded": round(cpu["threaded"]["value"] / ns_cell, 1)}
    line.update(extras)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
