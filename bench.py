"""Benchmark: SSP-RK3 steps of the explicit Euler residual on one MI355X.

Workload (the manufactured-solution tetrahedra, the configuration the metric is quoted on that
fits one GPU): the unit cube in n^3 cubes of 6 Kuhn tetrahedra (n = 128: 12,582,912 cells),
interior nodes jittered by +-10 % of the spacing (seed 0), cells renumbered along a Morton curve; a smooth manufactured
Euler solution (rho, u, v, w, p sinusoids of amplitude 0.1 about rho = 1, p = 1/1.4, Mach 0.3)
with its volume source and exact `state` boundary faces; second-order HLLC with face2
least-squares gradients and the Venkatakrishnan limiter (K = 5), fp32+acc64, SSP-RK3 with the
global CFL 0.5 time step. One step = one SSP-RK3 step (three residuals, the norms of R(u), the
spectral radius and time step, validity checks).

Prints one JSON line: ns per cell per step (lower is better), the roofline of the dominant kernel
(algorithmic bytes per launch over its event-timed duration, against the 8 TB/s HBM peak), and a
CPU baseline: the oracle (this repo's C++ restatement of the reference's CPU loops) timed on a
bounded sample on this host, serial and on every host thread the process may use.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(n: int, policy: str, flux: str, order: str = "morton"):
    import numpy as np

    from autan_amd import mesh as meshes
    from autan_amd import mms
    from autan_amd.stepper import freestream_scheme

    t0 = time.time()
    m = meshes.kuhn_tets(n, 0.1, 0, order=order)
    kinds = np.zeros(m.n_boundary, dtype=np.int32)
    meshes.set_boundary_kinds(m, kinds)
    log(f"mesh {n}^3 x 6 tets: {m.n_cells} cells, {m.n_faces} faces, {m.n_stencil} stencil entries ({time.time() - t0:.1f} s)")
    rho, vel, p = mms.freestream()
    sc = freestream_scheme(rho, vel, p, flux=flux, second_order=True, limiter=True)
    sc.venkat_k = 5.0
    x = m.centroids()
    u = mms.conserved(mms.primitive(x)).reshape(-1)
    src = mms.source(x).reshape(-1)
    bc = mms.primitive(m.face_centroids()[m.n_interior :])
    return m, sc, u, kinds.astype(np.float64), bc, src


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


def cpu_baseline(policy: str, flux: str, n: int, steps: int) -> dict:
    """The oracle (the reference CPU path's semantics: serial gather loops) on a bounded sample,
    then the same loops on every host thread this process may use."""
    import numpy as np

    from oracle import oracle as orc
    from oracle.oracle import Oracle

    m, sc, u0, kinds, bc, src = build_case(n, policy, flux)
    ora = Oracle(m, sc, policy, kinds, bc, src)
    timings = {}
    for threads in (1, host_threads()):
        orc.set_threads(threads)
        u = u0.copy()
        ora.step(u.copy(), cfl=0.5)  # warm
        t0 = time.perf_counter()
        for _ in range(steps):
            u, norms, h, inv = ora.step(u, cfl=0.5)
        timings[threads] = time.perf_counter() - t0
        assert inv is None and np.all(np.isfinite(norms))
    orc.set_threads(1)
    t1 = timings[1]
    tn_threads = max(timings)
    return {
        "value": round(t1 / steps / m.n_cells * 1e9, 2),
        "unit": "ns/cell/iter",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} SSP-RK3 steps of the same workload at {n}^3 x 6 = {m.n_cells} tets, serial oracle "
        f"(gather loops), {t1:.1f} s",
        "threaded": {
            "value": round(timings[tn_threads] / steps / m.n_cells * 1e9, 2),
            "cores": tn_threads,
            "cpu": cpu_model(),
            "seconds": round(timings[tn_threads], 2),
        },
    }


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cubes per side (6 n^3 tetrahedra)")
    ap.add_argument("--policy", default="fp32+acc64", choices=["fp32+acc64", "fp64"])
    ap.add_argument("--flux", default="hllc", choices=["roe", "hllc", "ausm+up"])
    ap.add_argument("--cpu-n", type=int, default=48, help="cubes per side of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--order", default="morton", choices=["morton", "rcm"], help="cell renumbering of the set-up")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has one, and this build has no multi-GPU path "
              "(the sharding design is in DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np

    from autan_amd import traffic
    from autan_amd.stepper import DeviceStepper

    m, sc, u, kinds, bc, src = build_case(args.n, args.policy, args.flux, args.order)
    dev = DeviceStepper(m, sc, args.policy, kinds, bc, src)
    dev.upload(u)
    log("uploaded; warming up")
    for _ in range(args.warmup):
        r = dev.step(cfl=0.5)
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = dev.step(cfl=0.5)
    elapsed = time.perf_counter() - t0
    if r.invalid is not None or not np.all(np.isfinite(r.residuals)):
        raise RuntimeError(f"the run diverged: invalid cell {r.invalid}, norms {r.residuals}")
    ms_step = elapsed / args.steps * 1e3
    ns_cell = ms_step * 1e6 / m.n_cells
    log(f"{args.steps} steps: {ms_step:.3f} ms/step, {ns_cell:.3f} ns/cell/step, norms {r.residuals}")

    prof = dev.profile(steps=5, cfl=0.5)
    counts = traffic.Counts.of(m, source=True, limiter=True, second_order=True)
    kb = traffic.kernel_bytes(counts, args.policy)
    per_step = {k: v["ms"] * v["launches"] for k, v in prof["kernels"].items()}
    dominant = max(per_step, key=per_step.get)
    ms_launch = prof["kernels"][dominant]["ms"]
    achieved = kb[dominant] / (ms_launch * 1e-3) / 1e9
    kernels = {
        k: {
            "ms": round(v["ms"], 4),
            "launches": v["launches"],
            "GB/s": round(kb[k] / (v["ms"] * 1e-3) / 1e9, 1) if k in kb and v["ms"] > 0 else None,
        }
        for k, v in prof["kernels"].items()
    }
    step_b = traffic.step_bytes(counts, args.policy)
    ref_b = traffic.reference_step_bytes(counts, args.policy)
    peak = stream_peak()
    log(f"profile: {json.dumps(kernels)}")
    log(f"STREAM: {json.dumps(peak)}")
    traffic_file = ROOT / "profiles" / f"traffic_{dominant}_tets{args.n}_{args.policy}.json"
    measured = json.loads(traffic_file.read_text())["bytes_per_launch"] if traffic_file.exists() else None

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, one thread)")
        cpu = cpu_baseline(args.policy, args.flux, args.cpu_n, args.cpu_steps)

    line = {
        "metric": "ns/cell/iter",
        "value": round(ns_cell, 4),
        "unit": "ns/cell/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc" if args.policy == "fp32+acc64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"Kuhn tetrahedra {args.n}^3 x 6 (jitter 10 %, seed 0), manufactured solution, "
            f"second-order {args.flux.upper()}, face2 LSQ + Venkatakrishnan K=5, {args.policy}, SSP-RK3 CFL 0.5, "
            f"{args.order} cell order",
            "cells": m.n_cells,
            "faces": m.n_faces,
            "stencil_entries": m.n_stencil,
        },
        "cells_iter_per_s": round(m.n_cells / (ms_step * 1e-3), 1),
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured,
            "algorithmic_bytes_per_launch": kb[dominant],
            "ms_per_launch": round(ms_launch, 4),
        },
        "step": {
            "event_ms": round(prof["ms_per_step"], 4),
            "algorithmic_bytes": step_b,
            "achieved_GBs": round(step_b / (ms_step * 1e-3) / 1e9, 1),
            "frac_of_peak": round(step_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "reference_B_step": ref_b,
            "reference_B_step_frac": round(ref_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "stream_GBs": {k: round(v, 1) for k, v in peak.items()},
            "kernels": kernels,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
