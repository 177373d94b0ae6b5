"""The C++ adapter include/mxeuler.hpp (the reference's DeviceStepper<P> and make_device_stepper<P>,
physics/include/autan/physics/euler_device.hpp:22-54) through its test program
tests/cpp/adapter_check.cpp: refused input raises std::invalid_argument before any device work, a
device failure std::runtime_error with E1101 (physics/src/device_stepper.cuh:179-193), and on a GPU
the adapter's calls give the bits of the same calls made through the Python binding."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cases
from autan_amd.stepper import POLICIES, DeviceStepper

ROOT = Path(__file__).resolve().parent.parent
PROGRAM = ROOT / "tests" / "_build" / "adapter_check"
CFL, DT = 0.5, 1e-4


def program() -> Path:
    if not PROGRAM.exists():
        raise RuntimeError(f"{PROGRAM} is not built: run autan_amd/build.py (or __graft_entry__.build())")
    return PROGRAM


def write_case(d: Path, case: cases.Case, policy: str, bc_kind=None) -> None:
    """The arrays DeviceStepper hands to mxeuler_create, as raw little-endian files."""
    m = case.mesh
    sc = case.scheme.native(m.all_planar)
    lines = [
        f"policy {POLICIES[policy]}",
        f"flux {sc.flux}",
        f"second_order {sc.second_order}",
        f"limiter {sc.limiter}",
        f"planar {sc.planar}",
        f"all_planar {int(m.all_planar)}",
        f"cfl {CFL!r}",
        f"dt {DT!r}",
    ]
    for k in ("gamma", "gas_constant", "mach_ref", "venkat_k", "length"):
        lines.append(f"{k} {getattr(sc, k)!r}")
    lines.append("scale " + " ".join(repr(v) for v in sc.scale))
    lines.append("reference " + " ".join(repr(v) for v in sc.reference))
    (d / "case.txt").write_text("\n".join(lines) + "\n")
    arrays = {
        "owner.i32": (m.owner, np.int32),
        "neighbour.i32": (m.neighbour, np.int32),
        "cell_face_offsets.i64": (m.cell_face_offsets, np.int64),
        "cell_faces.i32": (m.cell_faces, np.int32),
        "face_normal.f64": (m.face_normal, np.float64),
        "face_area.f64": (m.face_area, np.float64),
        "r_left.f64": (m.r_left, np.float64),
        "r_right.f64": (m.r_right, np.float64),
        "face_moment.f64": (m.face_moment, np.float64),
        "cell_volume.f64": (m.cell_volume, np.float64),
        "stencil_offsets.i64": (m.stencil_offsets, np.int64),
        "stencil.i32": (m.stencil, np.int32),
        "stencil_weights.f64": (m.stencil_weights, np.float64),
        "bc_kind.f64": (case.bc_kind if bc_kind is None else bc_kind, np.float64),
        "bc.f64": (case.bc, np.float64),
        "u.f64": (case.u, np.float64),
    }
    if case.source is not None:
        arrays["source.f64"] = (case.source, np.float64)
    for name, (a, dt) in arrays.items():
        a = np.asarray(a)
        assert a.dtype == dt or name.startswith(("bc", "u.", "source")), (name, a.dtype)
        np.ascontiguousarray(a, dtype=dt).tofile(d / name)


def run(d: Path) -> subprocess.CompletedProcess:
    return subprocess.run([str(program()), str(d)], capture_output=True, text=True, timeout=300)


def test_refused_input_raises_invalid_argument(tmp_path):
    """The library's own check (owner out of range) and the adapter's (boundary arrays one short)
    both arrive as std::invalid_argument, before any device is touched."""
    case = cases.fixture_case("hex", "roe")
    case.mesh.arrays["owner"] = case.mesh.owner.copy()
    case.mesh.arrays["owner"][3] = 10**6
    write_case(tmp_path, case, "fp64")
    p = run(tmp_path)
    assert p.returncode == 3, p.stdout + p.stderr
    assert p.stdout.startswith("invalid_argument: owner out of range"), p.stdout

    case = cases.fixture_case("hex", "roe")
    write_case(tmp_path, case, "fp32+acc64", bc_kind=case.bc_kind[:-1])
    p = run(tmp_path)
    assert p.returncode == 3, p.stdout + p.stderr
    assert "boundary arrays" in p.stdout


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible: the gpu test runs this case")
def test_without_a_device_the_factory_raises_runtime_error(tmp_path):
    """Valid input on a host without a GPU: std::runtime_error carrying E1101, as the reference's
    stepper reports a device failure."""
    case = cases.tets_case(3)
    write_case(tmp_path, case, "fp32+acc64")
    p = run(tmp_path)
    assert p.returncode == 4, p.stdout + p.stderr
    assert p.stdout.startswith("runtime_error: E1101"), p.stdout


def _read_results(path: Path) -> dict:
    out = {}
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 8:
            out[f[0]] = (np.array([float(x) for x in f[1:6]]), float(f[6]), int(f[7]))
        else:
            out[f[0]] = f[1]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,policy", [("tets", "fp32+acc64"), ("poly", "fp64"), ("poly", "fp32+acc64")])
def test_the_adapter_gives_the_bindings_bits(tmp_path, name, policy):
    """The same calls through the C++ adapter (a separate process) and through the Python binding:
    residual, evaluate, global, fixed and local steps, a first-order step, the final state."""
    case = cases.tets_case(6) if name == "tets" else cases.fixture_case("poly", "hllc")
    write_case(tmp_path, case, policy)
    print(f"adapter_check {name} {policy}", flush=True)
    p = run(tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    got = _read_results(tmp_path / "results.txt")
    res_cpp = np.fromfile(tmp_path / "res.f64")
    u_cpp = np.fromfile(tmp_path / "u_out.f64")

    dev = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    dev.upload(case.u)
    res = dev.residual(case.u).reshape(-1)
    want = {"evaluate": dev.evaluate(), "step_global": dev.step(CFL), "step_dt": dev.step(CFL, dt=DT)}
    dev.set_second_order(False)
    want["step_local_first_order"] = dev.step(CFL, local=True)
    dev.set_second_order(True)
    want["step_local"] = dev.step(CFL, local=True)
    u = dev.download().reshape(-1)
    dev.close()

    assert np.array_equal(res_cpp.view(np.uint64), res.view(np.uint64))
    for k, r in want.items():
        rr, ts, inv = got[k]
        assert np.array_equal(rr.view(np.uint64), np.asarray(r.residuals, dtype=np.float64).view(np.uint64)), k
        assert ts == r.time_step or k == "evaluate", k
        assert inv == (-1 if r.invalid is None else r.invalid), k
    assert got["step_dt"][1] == DT
    assert got["moved_from"] == "refused"
    assert np.array_equal(u_cpp.view(np.uint64), u.view(np.uint64))
    assert not np.array_equal(u, np.asarray(case.u).reshape(-1))  # the steps moved the state
