"""history.jsonl lines (autan_amd/history.py) against the reference runtime's JSON library:
golden numbers and lines printed by nlohmann::ordered_json's dump() (tests/golden/
make_history_golden.py through oracle/_ref/json_dump), including doubles whose Grisu2 digits are
not the shortest; the library itself on fresh values where oracle/_ref is built; the line rules of
docs/contracts/results.md (History) as py/autan/results.py:911-959 reads them."""

from __future__ import annotations

import json
import math
import struct
import subprocess
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from autan_amd import history

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
DUMP = ROOT / "oracle" / "_ref" / "json_dump"


def from_bits(h: str) -> float:
    return struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0]


def bits(x: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}"


def golden_numbers():
    for line in (GOLDEN / "history_numbers.txt").read_text().splitlines():
        h, text = line.split(" ")
        yield from_bits(h), text


def test_numbers_are_the_runtime_librarys_text():
    rows = list(golden_numbers())
    assert len(rows) > 2000
    wrong = [(x, t, history.format_double(x)) for x, t in rows if history.format_double(x) != t]
    assert not wrong, wrong[:5]


def test_the_fixture_holds_digits_that_are_not_the_shortest():
    """At least some golden texts differ from the shortest-nearest digits, so a formatter built on
    Python's repr would fail the test above."""
    def digits(text):
        return text.lstrip("-").split("e")[0].replace(".", "").strip("0")

    differ = [x for x, t in golden_numbers() if math.isfinite(x) and x != 0 and digits(t) != digits(repr(abs(x)))]
    assert len(differ) >= 20


def _parse_request(req: str):
    f = req.split()
    kind, it, second, res = f[0], int(f[1]), f[2], [from_bits(h) for h in f[3:8]]
    if kind == "s":
        return history.record(it, res, provisional=second == "1")
    return history.record(it, res, time=from_bits(second))


def test_lines_are_the_runtimes_lines():
    rows = [line.split("\t") for line in (GOLDEN / "history_lines.txt").read_text().splitlines()]
    assert len(rows) == 40
    for req, want in rows:
        assert _parse_request(req) == want, req


@pytest.mark.skipif(not DUMP.exists(), reason="oracle/_ref/json_dump is built only where the reference tree is")
def test_against_the_runtime_library_on_fresh_values():
    rng = np.random.default_rng(99)
    raw = rng.integers(0, 2**63 - 1, 30000, dtype=np.int64).astype(np.uint64).view(np.float64)
    vals = [float(v) for v in raw[np.isfinite(raw)]] + [float(v) for v in 10 ** rng.uniform(-16, 4, 5000)]
    out = subprocess.run([str(DUMP)], input="".join(f"n {bits(v)}\n" for v in vals), capture_output=True,
                         text=True, check=True).stdout.splitlines()
    wrong = [(v, o) for v, o in zip(vals, out) if history.format_double(v) != o]
    assert not wrong, wrong[:5]


def test_every_text_reads_back_as_the_same_double():
    """Round trip: the text parses to the double it was written from (the property Grisu2 keeps
    even where its digits are not the shortest)."""
    rng = np.random.default_rng(5)
    raw = rng.integers(0, 2**64 - 1, 20000, dtype=np.uint64).view(np.float64)
    for x in raw[np.isfinite(raw)]:
        t = history.format_double(float(x))
        assert float(t) == x and (math.copysign(1.0, float(t)) == math.copysign(1.0, x)), (x, t)
    assert history.format_double(float("nan")) == "null"
    assert history.format_double(float("inf")) == "null"


def test_layout_switches_where_the_library_switches():
    """The decimal point within 15 digits left and 4 zeros right of the digits, else an exponent of
    at least two digits; whole numbers keep a trailing .0."""
    assert history.format_double(1e14) == "100000000000000.0"
    assert history.format_double(999999999999999.0) == "999999999999999.0"
    assert history.format_double(1e15) == "1e+15"
    assert history.format_double(0.0001) == "0.0001"
    assert history.format_double(1e-5) == "1e-05"
    assert history.format_double(1.5e-5) == "1.5e-05"
    assert history.format_double(-2.5) == "-2.5"
    assert history.format_double(-0.0) == "-0.0"
    assert history.format_double(5e-324) == "5e-324"
    assert history.format_double(1.7976931348623157e308) == "1.7976931348623157e+308"


def _problems(records, steady: bool) -> list[str]:
    """The line rules results.md states for history.jsonl (keys, increasing integer iterations,
    time null exactly for steady runs, the five equations in order, finite non-negative values)."""
    found, previous = [], 0
    keys = {"iteration", "time", "residuals", "monitors"} | ({"provisional"} if steady else set())
    for r in records:
        if set(r) != keys:
            found.append(f"keys {sorted(r)}")
            continue
        if not isinstance(r["iteration"], int) or r["iteration"] <= previous:
            found.append("iterations")
        previous = r["iteration"]
        if steady != (r["time"] is None):
            found.append("time")
        if list(r["residuals"]) != list(history.EQUATIONS):
            found.append("equations")
        elif not all(isinstance(v, float) and math.isfinite(v) and v >= 0 for v in r["residuals"].values()):
            found.append("residual values")
        if steady and not isinstance(r["provisional"], bool):
            found.append("provisional")
    return found


def test_steady_lines_of_a_run(tmp_path):
    rng = np.random.default_rng(1)
    n = 250
    res = SimpleNamespace(history=10 ** rng.uniform(-12, 0, (n, 5)), provisional=np.arange(n) < 37)
    lines = history.steady_lines(res, first_iteration=101)
    path = tmp_path / "history.jsonl"
    history.write(path, lines[:100])
    history.write(path, lines[100:], append=True)
    with path.open("ab") as f:
        f.write(b'{"iteration":351,"ti')  # a live run's partial last line is skipped
    records = history.read(path)
    assert len(records) == n
    assert not _problems(records, steady=True)
    assert [r["iteration"] for r in records] == list(range(101, 101 + n))
    assert sum(r["provisional"] for r in records) == 37 and all(r["provisional"] for r in records[:37])
    got = np.array([[r["residuals"][k] for k in history.EQUATIONS] for r in records])
    assert np.array_equal(got.view(np.uint64), res.history.view(np.uint64))
    assert path.read_bytes().split(b"\n")[0].startswith(b'{"iteration":101,"time":null,"residuals":{"density":')


def test_unsteady_lines_accumulate_time_in_fp64():
    steps = [SimpleNamespace(residuals=[1e-3, 2e-4, 3e-4, 4e-4, 5e-3], time_step=0.1) for _ in range(10)]
    records = [json.loads(x) for x in history.unsteady_lines(steps, start_time=1.0)]
    assert not _problems(records, steady=False)
    t = 1.0
    for r in records:
        t = t + 0.1
        assert r["time"] == t
    assert "provisional" not in records[0]


def test_the_forces_monitor():
    """Keys in the runtime's order; at zero incidence the coefficients are the force components
    over q A (drag x, side y, lift z); at 90 degrees of incidence lift is -x and drag z."""
    f, m = (3.0, -1.5, 6.0), (0.5, 2.0, -4.0)
    assert list(history.forces(f, m)) == ["fx", "fy", "fz", "mx", "my", "mz"]
    r = history.forces(f, m, ref_area=2.0, ref_length=0.5, density=1.2, speed=10.0)
    assert list(r) == ["fx", "fy", "fz", "mx", "my", "mz", "cd", "cs", "cl", "cmx", "cmy", "cmz"]
    qa = 0.5 * 1.2 * 100.0 * 2.0
    assert (r["cd"], r["cs"], r["cl"]) == (3.0 / qa, -1.5 / qa, 6.0 / qa)
    assert (r["cmx"], r["cmy"], r["cmz"]) == (0.5 / (qa * 0.5), 2.0 / (qa * 0.5), -4.0 / (qa * 0.5))
    r = history.forces(f, m, ref_area=2.0, density=1.2, speed=10.0, aoa=math.pi / 2)
    assert "cmx" not in r
    assert abs(r["cl"] + 3.0 / qa) < 1e-15 and abs(r["cd"] - 6.0 / qa) < 1e-15 and abs(r["cs"] + 1.5 / qa) < 1e-15
    line = history.record(5, [1e-3] * 5, provisional=False, monitors={"forces": r})
    assert '"monitors":{"forces":{"fx":3.0,"fy":-1.5,"fz":6.0,"mx":0.5,"my":2.0,"mz":-4.0,"cd":' in line


def test_monitors_and_refusals():
    line = history.record(3, [1.0] * 5, provisional=False,
                          monitors={"forces": {"fx": 1.25, "fy": 0.0, "fz": -3e-7}})
    assert line.endswith('"monitors":{"forces":{"fx":1.25,"fy":0.0,"fz":-3e-07}},"provisional":false}')
    with pytest.raises(ValueError):
        history.record(1, [1.0] * 4, provisional=False)
    with pytest.raises(ValueError):
        history.record(1, [1.0] * 5)
    with pytest.raises(ValueError):
        history.record(1, [1.0] * 5, time=0.5, provisional=True)


def test_an_oracle_steady_runs_history():
    """The oracle's steady run (first-order start, then second order) written as history lines:
    one per completed iteration, its residuals' bits, the first-order start's lines provisional."""
    import cases
    from autan_amd import steady
    from oracle.oracle import Oracle

    case = cases.tets_case(3)
    ora = Oracle(case.mesh, case.scheme, "fp64", case.bc_kind, case.bc, case.source)
    run = steady.solve_with(cases.OracleTarget(ora, case.u), max_iterations=150, residual_orders=8.0)
    records = [json.loads(x) for x in history.steady_lines(run)]
    assert not _problems(records, steady=True)
    assert len(records) == run.iterations and records[-1]["iteration"] == run.iterations
    assert sum(r["provisional"] for r in records) == run.first_order_steps > 0
    got = np.array([[r["residuals"][k] for k in history.EQUATIONS] for r in records])
    assert np.array_equal(got.view(np.uint64), run.history.view(np.uint64))


@pytest.mark.gpu
def test_a_device_steady_runs_history_is_the_oracle_runs():
    """The device's steady run (first-order start, fp32+acc64 level rule) written as history.jsonl
    is the oracle run's text line for line; one line per completed iteration, the first-order
    start's lines provisional."""
    import cases
    from autan_amd import steady
    from autan_amd.stepper import DeviceStepper
    from oracle.oracle import Oracle

    case = cases.tets_case(4)
    policy = "fp32+acc64"
    dev = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    ora = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    dev.upload(case.u)
    ref = np.array(case.scheme.reference)
    speed = math.sqrt(ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2) / ref[0]
    p = 0.4 * (ref[4] - 0.5 * ref[0] * speed**2)
    kw = dict(max_iterations=240, residual_orders=8.0, level_rule=True, scales=dev.residual_scales(ref[0], speed, p))
    print("steady run: tets 4, fp32+acc64, 240 iterations", flush=True)
    got = steady.solve(dev, **kw)
    want = steady.solve_with(cases.OracleTarget(ora, case.u), **kw)
    dev.close()
    lines = history.steady_lines(got)
    assert lines == history.steady_lines(want)
    records = [json.loads(x) for x in lines]
    assert not _problems(records, steady=True)
    assert len(records) == got.iterations and records[-1]["iteration"] == got.iterations
    assert sum(r["provisional"] for r in records) == got.first_order_steps > 0
