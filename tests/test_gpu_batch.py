"""GPU parity of batches: several bump channels of different size, bump height, Mach number and
flow angle stepped together by one stepper give, case by case, the bits of each case stepped
alone."""

from __future__ import annotations

import numpy as np
import pytest

from autan_amd import batch, steady, workloads
from autan_amd.stepper import DeviceStepper

pytestmark = pytest.mark.gpu

SPECS = [  # (nx, ny, nz, h, mach, alpha)
    (24, 8, 4, 0.05, 0.4, 2.0),
    (32, 12, 8, 0.1, 0.6, -3.0),
    (16, 6, 4, 0.02, 0.75, 0.5),
]


def make_cases():
    out = []
    for nx, ny, nz, h, mach, alpha in SPECS:
        w = workloads.bump(nx, ny, nz, "fp32+acc64", "hllc", h=h, mach=mach, alpha_deg=alpha, prism_layers=2)
        w.scheme.planar = False  # the batch runs the warped-face loops of its non-planar cases
        out.append(batch.Case(w.mesh, w.scheme, w.u, w.bc_kind, w.bc))
    return out


def solo(c, policy):
    return DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_batch_steps_are_each_cases_steps(policy):
    cs = make_cases()
    b = batch.BatchStepper(cs, policy)
    b.upload()
    alone = [solo(c, policy) for c in cs]
    for d, c in zip(alone, cs):
        d.upload(c.u)
    for _ in range(3):
        r, h, inv = b.step(cfl=0.8, local=True)
        for j, d in enumerate(alone):
            s = d.step(cfl=0.8, local=True)
            np.testing.assert_array_equal(r[j], s.residuals)
            assert inv[j] is None and s.invalid is None
    for j, (got, d) in enumerate(zip(b.download(), alone)):
        np.testing.assert_array_equal(got, d.download())
    # global time steps: each case its own least step
    r, h, inv = b.step(cfl=0.5, local=False)
    for j, d in enumerate(alone):
        s = d.step(cfl=0.5)
        np.testing.assert_array_equal(r[j], s.residuals)
        assert h[j] == s.time_step


def test_inactive_cases_keep_their_state():
    cs = make_cases()
    b = batch.BatchStepper(cs, "fp32+acc64")
    b.upload()
    b.step(cfl=0.8, local=True)
    frozen = b.download()[1]
    b.set_flags([3, 2, 3])
    for _ in range(3):
        b.step(cfl=0.8, local=True)
    after = b.download()
    np.testing.assert_array_equal(after[1], frozen)
    assert not np.array_equal(after[0], cs[0].u)


@pytest.mark.parametrize("graph", [0, 1])
def test_a_case_stepped_again_continues_from_its_frozen_state(graph):
    """A case frozen over an odd number of steps (the buffers the others step through change
    parity) and then stepped again gives the bits of the same case stepped alone without the pause."""
    cs = make_cases()
    b = batch.BatchStepper(cs, "fp32+acc64")
    b.tune("graph", graph)
    b.upload()
    alone = solo(cs[1], "fp32+acc64")
    alone.upload(cs[1].u)
    b.step(cfl=0.8, local=True)
    alone.step(cfl=0.8, local=True)
    b.set_flags([3, 2, 3])
    for _ in range(3):
        b.step(cfl=0.8, local=True)
    b.set_flags([3, 3, 3])
    for _ in range(2):
        r, _, inv = b.step(cfl=0.8, local=True)
        s = alone.step(cfl=0.8, local=True)
        np.testing.assert_array_equal(r[1], s.residuals)
        assert inv[1] is None
    np.testing.assert_array_equal(b.download()[1], alone.download())


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("start", ["second", "first"])
def test_batch_steady_matches_each_solo_run(start, graph):
    cs = make_cases()
    policy = "fp32+acc64"
    b = batch.BatchStepper(cs, policy)
    b.tune("graph", graph)
    b.upload()
    kw = dict(max_iterations=230, residual_orders=8.0, start_order=start)
    got = b.steady(**kw)
    states = b.download()
    for j, c in enumerate(cs):
        d = solo(c, policy)
        d.upload(c.u)
        want = steady.solve(d, **kw)
        assert got[j].iterations == want.iterations and got[j].status == want.status
        assert got[j].first_order_steps == want.first_order_steps and got[j].first_order_end == want.first_order_end
        np.testing.assert_array_equal(got[j].history, want.history)
        np.testing.assert_array_equal(states[j], d.download())


def test_autotune_restores_the_state_and_keeps_the_bits():
    cs = make_cases()
    b = batch.BatchStepper(cs, "fp32+acc64")
    b.upload()
    report = b.autotune(steps=2)
    assert report.startswith("best face_block=")
    for got, c in zip(b.download(), cs):
        np.testing.assert_array_equal(got, c.u)
    r, h, inv = b.step(cfl=0.8, local=True)
    for j, c in enumerate(cs):
        d = solo(c, "fp32+acc64")
        d.upload(c.u)
        np.testing.assert_array_equal(r[j], d.step(cfl=0.8, local=True).residuals)
