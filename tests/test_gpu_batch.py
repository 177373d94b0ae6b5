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



def sweep_cases(policy, inflow="farfield"):
    """Cases of the sweep's own generator and builder (seed 1: the drawn heights, Mach numbers and
    angles; two prism rows, the warped-face loops), on meshes small enough for the oracle. With
    inflow "state", the channel's farfield faces take the freestream as `state` data, so that no
    face calls exp/log."""
    from autan_amd import mesh as meshes
    from autan_amd import sweep

    specs = sweep.specs(4, seed=1, cells=(2e3, 3e4))
    cs = [sweep.build(s, policy) for s in specs]
    if inflow == "state":
        for c in cs:
            kinds = np.where(c.bc_kind == meshes.FARFIELD, float(meshes.STATE), c.bc_kind)
            c.bc_kind = kinds
            meshes.set_boundary_kinds(c.mesh, kinds.astype(np.int32))
    return specs, cs


@pytest.mark.parametrize("inflow", ["state", "farfield"])
@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_batch_steps_match_the_oracle(policy, inflow):
    """The batch path against the CPU oracle, case by case: states, norms and time steps of local
    and global SSP-RK3 steps. Bit for bit where no face calls exp/log; with farfield faces (exp/log
    from glibc on the host, from the device library on the GPU) within the rounding those leave,
    each step continuing from the device's state."""
    from oracle import oracle as orc
    from oracle.oracle import Oracle

    orc.set_threads(8)
    try:
        _, cs = sweep_cases(policy, inflow)
        b = batch.BatchStepper(cs, policy)
        b.upload()
        oracles = [Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc) for c in cs]
        states = [c.u.copy() for c in cs]
        exact = inflow == "state"
        for it in range(4):
            local = it < 3
            cfl = 0.8 if local else 0.5
            r, h, inv = b.step(cfl=cfl, local=local)
            got = b.download()
            for j, ora in enumerate(oracles):
                u, norms, dt, bad = ora.step(states[j], cfl=cfl, local=local)
                assert inv[j] is None and bad is None
                if exact:
                    np.testing.assert_array_equal(got[j], u)
                    np.testing.assert_array_equal(r[j], norms)
                    if not local:
                        assert h[j] == dt
                else:
                    # The state's own magnitudes: density, momentum (one scale for its three
                    # components: the spanwise one is near zero in these channels, while a
                    # rounding difference in a boundary pressure moves every component alike),
                    # total energy. 64 units of the compute type of that scale.
                    tol = 1e-12 if policy == "fp64" else 64 * float(np.finfo(np.float32).eps)
                    a = np.abs(u).reshape(5, -1).max(axis=1)
                    scale = np.array([a[0], a[1:4].max(), a[1:4].max(), a[1:4].max(), a[4]])[:, None]
                    diff = np.abs(got[j].reshape(5, -1) - u.reshape(5, -1))
                    assert np.all(diff <= tol * scale), (it, j, (diff / scale).max(axis=1))
                    np.testing.assert_allclose(r[j], norms, rtol=1e-9 if policy == "fp64" else 1e-4)
                    if not local:
                        assert h[j] == pytest.approx(dt, rel=1e-9 if policy == "fp64" else 1e-5)
            states = [g.copy() for g in got]
    finally:
        orc.set_threads(1)


@pytest.mark.parametrize("inflow", ["state", "farfield"])
def test_batch_steady_run_matches_the_oracle(inflow):
    """A short steady run of every sweep case in one batch (first-order start, local steps, the
    convergence rule with each case's round-off floors) and the same run of the oracle: the same
    iteration counts, first-order start and status; the residual histories bit for bit where no
    face calls exp/log."""
    from autan_amd import sweep
    from oracle import oracle as orc
    from oracle.oracle import Oracle

    import cases as tc

    orc.set_threads(8)
    try:
        policy = "fp32+acc64"
        _, cs = sweep_cases(policy, inflow)
        b = batch.BatchStepper(cs, policy)
        b.tune("graph", 1)
        b.upload()
        scales = np.array([sweep.residual_scales(c) for c in cs])
        kw = dict(max_iterations=240, residual_orders=8.0, start_order="first")
        got = b.steady(scales=scales, **kw)
        states = b.download()
        for j, c in enumerate(cs):
            target = tc.OracleTarget(Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc), c.u)
            want = steady.solve_with(target, level_rule=True, scales=scales[j], **kw)
            assert got[j].iterations == want.iterations and got[j].status == want.status, j
            assert got[j].first_order_steps == want.first_order_steps
            assert got[j].first_order_end == want.first_order_end
            if inflow == "state":
                np.testing.assert_array_equal(got[j].history, want.history)
                np.testing.assert_array_equal(states[j], target.u)
            else:
                np.testing.assert_allclose(got[j].history, want.history, rtol=1e-3, atol=0)
    finally:
        orc.set_threads(1)
