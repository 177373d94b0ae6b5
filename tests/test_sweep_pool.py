"""The implicit sweep's case pool (sweep.solve_implicit_cases) on the host, with stand-ins for the
device stepper and its solve: every case solved once and reported in case order whatever order the
solves finish in, `finish` called in case order in the calling thread, the chunks built before they
are solved, every stepper closed once, and a failing solve raised after the chunk's steppers are
closed."""

from __future__ import annotations

import random
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

import autan_amd.stepper
from autan_amd import steady, sweep


class FakeStepper:
    made: list = []
    lock = threading.Lock()

    def __init__(self, mesh, scheme, policy, bc_kind, bc):
        self.k = mesh.k
        self.closed = 0
        self.uploaded = None
        with FakeStepper.lock:
            FakeStepper.made.append(self)

    def upload(self, u):
        self.uploaded = u

    def close(self):
        self.closed += 1


@pytest.fixture
def pool(monkeypatch):
    FakeStepper.made = []
    live = {"built": 0, "max_built_unsolved": 0, "closed_while_solving": 0}
    lock = threading.Lock()

    def build(spec, policy, flux):
        with lock:
            live["built"] += 1
        return SimpleNamespace(mesh=SimpleNamespace(k=spec, n_cells=100 + spec), scheme=None, bc_kind=None, bc=None,
                               u=np.full(3, float(spec)))

    def solve(dev, max_iterations, residual_orders, scales, reference, start_order):
        with lock:
            live["max_built_unsolved"] = max(live["max_built_unsolved"], live["built"])
            live["built"] -= 1
        time.sleep(random.uniform(0.0, 0.01))
        if dev.k == failing.get("k"):
            raise RuntimeError(f"case {dev.k} failed")
        time.sleep(0.02)
        if dev.closed:
            with lock:
                live["closed_while_solving"] += 1
        return SimpleNamespace(iterations=dev.k + 1, status=0, uploaded=dev.uploaded, reference=reference)

    failing: dict = {}
    monkeypatch.setattr(sweep, "build", build)
    monkeypatch.setattr(sweep, "residual_scales", lambda c: np.ones(5))
    monkeypatch.setattr(autan_amd.stepper, "DeviceStepper", FakeStepper)
    monkeypatch.setattr(steady, "solve_implicit", solve)
    return live, failing


@pytest.mark.parametrize("workers,chunk", [(1, 64), (4, 64), (4, 5), (8, 3)])
def test_every_case_once_in_case_order(pool, workers, chunk):
    live, _ = pool
    specs = list(range(13))
    seen = []

    def finish(k, c, dev, r):
        seen.append((k, threading.current_thread() is threading.main_thread(), dev.closed))

    states = [np.full(3, 10.0 + k) for k in specs]
    refs = [np.full(5, 20.0 + k) for k in specs]
    out, wall = sweep.solve_implicit_cases(specs, "fp64", states=states, references=refs, workers=workers,
                                           chunk=chunk, finish=finish, log=lambda *a: None)
    assert [o["result"].iterations for o in out] == [k + 1 for k in specs]
    assert [o["cells"] for o in out] == [100 + k for k in specs]
    assert all(o["seconds"] >= 0.0 for o in out) and wall > 0.0
    # each case starts from its own state and is held to its own reference
    assert all(np.array_equal(o["result"].uploaded, states[k]) for k, o in enumerate(out))
    assert all(np.array_equal(o["result"].reference, refs[k]) for k, o in enumerate(out))
    # finish sees the cases in order, in the calling thread, before their steppers close
    assert seen == [(k, True, 0) for k in specs]
    assert sorted(d.k for d in FakeStepper.made) == specs
    assert all(d.closed == 1 for d in FakeStepper.made)
    # a chunk is built before any of it is solved, and no more than a chunk is built at once
    assert live["max_built_unsolved"] <= chunk
    assert live["closed_while_solving"] == 0


def test_a_failing_solve_raises_and_closes_the_chunk(pool):
    live, failing = pool
    failing["k"] = 6
    with pytest.raises(RuntimeError, match="case 6 failed"):
        sweep.solve_implicit_cases(list(range(10)), "fp64", workers=3, chunk=4, log=lambda *a: None)
    # the chunks that were built (0-3, 4-7) closed every stepper, none while its solve ran; the last
    # chunk never started
    assert sorted(d.k for d in FakeStepper.made) == list(range(8))
    assert all(d.closed == 1 for d in FakeStepper.made)
    assert live["closed_while_solving"] == 0
