"""The steady implicit path on the device against the CPU oracle, through the C ABI:

- the residual's Jacobian parts on warped meshes (the warp term's derivative included), bit for bit;
- one implicit solve and update (assembly of lambda / CFL I + dR/dU, the inverted diagonal blocks,
  GCR with exact inner products preconditioned by the symmetric multicolour sweeps, the admissible
  fraction): the new state, the solve's report and the colours, bit for bit;
- whole steady implicit runs (first-order start, switched evolution relaxation, the convergence
  rule): every reported residual, CFL number, GCR count and update fraction bit for bit where no face
  calls exp/log, and to 1e-8 on sweep channels, the transonic one the explicit method loses
  included."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd import steady, sweep
from autan_amd.stepper import DeviceStepper
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

STATE_ZONES = [meshes.STATE] * 6


def state_inflow(c):
    """The channel with its farfield faces turned into `state` faces holding the freestream: no
    face calls exp/log, so the device and the oracle agree bit for bit."""
    kinds = np.where(c.bc_kind == meshes.FARFIELD, float(meshes.STATE), c.bc_kind)
    c.bc_kind = kinds
    meshes.set_boundary_kinds(c.mesh, kinds.astype(np.int32))
    return c


def sweep_case(index: int, cells: float, policy: str, inflow: str = "farfield"):
    sp = dataclasses.replace(sweep.specs(64, seed=1)[index], lattice=sweep.lattice_for(cells))
    c = sweep.build(sp, policy)
    return state_inflow(c) if inflow == "state" else c


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("flux", ["roe", "hllc", "ausm+up"])
def test_warped_jacobian_parts_are_the_oracles(flux, policy):
    c = cases.fixture_case("poly", flux=flux, kinds_by_zone=STATE_ZONES)
    assert not c.mesh.all_planar
    print(f"warped jacobian parts poly {flux} {policy}", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc, c.source)
    got = dev.residual_and_jacobian(c.u)
    dev.close()
    want = orc.Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc, c.source).residual_and_jacobian(c.u)
    for k in ("res", "dwdu", "face_jac", "bnd_jac"):
        np.testing.assert_array_equal(np.asarray(got[k]).reshape(-1), want[k], err_msg=k)


ALL_KINDS = [meshes.FARFIELD, meshes.OUTLET, meshes.SLIP, meshes.WALL, meshes.INLET, meshes.STATE]


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("second,cfl,zones", [(False, 1000.0, "state"), (True, 50.0, "state"), (True, 1e6, "state"),
                                              (True, 200.0, "all")])
def test_one_implicit_solve_is_the_oracles(policy, second, cfl, zones):
    """At second order the Krylov operator carries the boundary faces' gradient term; with every
    boundary kind (farfield and inlet call exp/log) the device and the oracle agree to rounding."""
    kinds = STATE_ZONES if zones == "state" else ALL_KINDS
    c = cases.fixture_case("poly", flux="hllc", second_order=second, kinds_by_zone=kinds)
    print(f"implicit solve poly hllc second={second} cfl={cfl} {policy}", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc, c.source)
    dev.upload(c.u)
    norms, inv = dev.implicit_evaluate()
    rep = dev.implicit_solve(cfl)
    u_dev = dev.download()
    colours = dev.implicit_colours()
    dev.close()
    ora = orc.Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc, c.source)
    parts = ora.residual_and_jacobian(c.u)
    r = parts["res"].reshape(5, -1)
    want_norms = [np.sqrt(orc.exact_sum(r[k] * r[k]) / r.shape[1]) for k in range(5)]
    if zones == "state":
        np.testing.assert_array_equal(norms, want_norms)
    else:
        np.testing.assert_allclose(norms, want_norms, rtol=1e-9 if policy == "fp64" else 1e-4)
    assert inv is None
    u_ora, du, want, col = orc.implicit_solve(c.mesh, parts, ora.spectral_radius(), cfl, c.scheme.gamma, c.u,
                                              oracle=ora, second=second)
    np.testing.assert_array_equal(colours, col)
    assert not rep["failed"] and rep["alpha"] > 0
    if zones == "state":
        assert rep == want
        np.testing.assert_array_equal(u_dev, u_ora)
    else:
        assert rep["cycles"] == want["cycles"] and rep["alpha"] == want["alpha"]
        assert rep["residual_drop"] == pytest.approx(want["residual_drop"], rel=1e-6)
        tol = 1e-10 if policy == "fp64" else 1e-4
        np.testing.assert_allclose(u_dev, u_ora, rtol=0, atol=tol * np.abs(u_ora).max())


@pytest.mark.parametrize("policy,budget", [("fp64", 80), ("fp32+acc64", 300)])
def test_an_implicit_steady_run_is_the_oracles(policy, budget):
    """fp64 converges to 1e-8; fp32+acc64 levels off at its round-off floor (the fluxes in FP32)
    and converges by the plateau rule, which watches 100 iterations of the level after 100 more."""
    c = sweep_case(3, 1200, policy, inflow="state")
    print(f"implicit steady run sweep case 3 ({c.mesh.n_cells} cells) {policy}", flush=True)
    scales = sweep.residual_scales(c)
    dev = DeviceStepper(c.mesh, c.scheme, policy, c.bc_kind, c.bc)
    dev.upload(c.u)
    got = steady.solve_implicit(dev, max_iterations=budget, residual_orders=8.0, scales=scales)
    u_dev = dev.download()
    dev.close()
    orc.set_threads(8)
    try:
        ora = orc.Oracle(c.mesh, c.scheme, policy, c.bc_kind, c.bc)
        t = cases.OracleImplicitTarget(ora, c.u, c.mesh, c.scheme.gamma)
        want = steady.solve_implicit_with(t, max_iterations=budget, residual_orders=8.0, level_rule=policy == "fp32+acc64",
                                          scales=scales)
    finally:
        orc.set_threads(1)
    assert (got.status, got.iterations, got.first_order_steps, got.first_order_end) == \
        (want.status, want.iterations, want.first_order_steps, want.first_order_end)
    np.testing.assert_array_equal(got.history, want.history)
    np.testing.assert_array_equal(got.cfl, want.cfl)
    np.testing.assert_array_equal(got.cycles, want.cycles)
    np.testing.assert_array_equal(got.alpha, want.alpha)
    np.testing.assert_array_equal(u_dev, t.u)
    assert got.status == steady.CONVERGED
    if policy == "fp32+acc64":
        assert got.plateau_mask != 0 and got.iterations < budget


@pytest.mark.parametrize("index", [59, 0])
def test_implicit_runs_reach_the_threshold_on_sweep_channels(index):
    """fp64 to a 1e-8 relative residual from the freestream, farfield inflow: case 59 (Mach 0.77,
    bump 0.1045, flow angle -3.8 deg) is the channel whose explicit run diverges at its inflow
    corner; case 0 a subsonic one."""
    c = sweep_case(index, 6000, "fp64")
    print(f"implicit steady run sweep case {index} ({c.mesh.n_cells} cells) fp64", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, "fp64", c.bc_kind, c.bc)
    dev.upload(c.u)
    r = steady.solve_implicit(dev, max_iterations=800, residual_orders=8.0, scales=sweep.residual_scales(c))
    dev.close()
    assert r.status == steady.CONVERGED, (r.iterations, r.residuals / r.reference)
    assert np.all(r.alpha[r.alpha > 0] <= 1.0)


def test_cases_solved_together_give_each_case_its_own_bits():
    """Four sweep channels solved four at a time (one stepper, stream and host thread each) give
    every case the residuals, CFL numbers, update fractions and final state of the same case solved
    alone: the steppers share nothing but the device."""
    sp = [dataclasses.replace(s, lattice=sweep.lattice_for(5000)) for s in sweep.specs(64, seed=1)[:4]]
    print(f"implicit cases together: {len(sp)} channels of {sweep.lattice_for(5000)}", flush=True)
    runs = {}
    for workers in (4, 1):
        finals = {}

        def keep(k, c, dev, r):
            finals[k] = dev.download()

        solved, wall = sweep.solve_implicit_cases(sp, "fp64", max_iterations=40, workers=workers, finish=keep,
                                                  log=lambda *a: None)
        assert wall > 0
        runs[workers] = (solved, finals)
    for k in range(len(sp)):
        a, b = runs[4][0][k]["result"], runs[1][0][k]["result"]
        assert a.iterations == b.iterations > 0
        np.testing.assert_array_equal(a.history, b.history)
        np.testing.assert_array_equal(a.cfl, b.cfl)
        np.testing.assert_array_equal(a.alpha, b.alpha)
        np.testing.assert_array_equal(runs[4][1][k], runs[1][1][k])
