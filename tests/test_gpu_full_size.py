"""The headline configuration at its full size: the 7,915,215-cell hybrid bump channel (462 x 154 x
154 lattice, prism rows, merged polyhedra with warped faces), second-order HLLC with the
Venkatakrishnan limiter, the mesh and scheme the benchmark times, on a smooth perturbation of its
freestream (so that gradients and limiter act in every cell). What does not depend on size:

- with the farfield faces holding the freestream as `state` faces (no exp/log), the device's
  fp32+acc64 residual, one SSP-RK3 step on local time steps (state, exact norms, least time step)
  and the conservation sums equal the oracle's bit for bit, the oracle running on the host's
  threads;
- with the farfield faces, the fp32+acc64 residual lies within the reference's 500 eps32 f_k A_i of
  the fp64 one on the same state, conservation E stays within its round-off bound, and repeated
  runs give the same bits.
The manufactured-solution tetrahedra at their largest level (12.6M cells) too: the device residual
is the oracle's bit for bit; and the 64-case sweep in one batch: each case's steps are its solo
steps' bits."""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd import workloads
from autan_amd.stepper import DeviceStepper
from oracle import oracle as orc
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def channel():
    print("full-size bump channel 462 x 154 x 154: building the mesh", flush=True)
    w = workloads.bump(policy="fp32+acc64")
    x = w.mesh.centroids()
    wf = workloads.bump_freestream()
    s = np.sin(2 * math.pi * x[:, 0] / 3.0) * np.sin(2 * math.pi * x[:, 1]) * np.cos(2 * math.pi * x[:, 2])
    u = cases.conserved(wf[0] * (1 + 0.02 * s), wf[1] * (1 + 0.05 * s), 0.01 * s, 0.01 * np.roll(s, 7),
                        wf[4] * (1 + 0.03 * s)).reshape(-1)
    return w, u


@pytest.fixture(scope="module")
def stepper(channel):
    """The fp32+acc64 device stepper of the channel with its farfield faces (one for the module:
    creating one uploads the 8M-cell mesh)."""
    w, _ = channel
    dev = DeviceStepper(w.mesh, w.scheme, "fp32+acc64", w.bc_kind, w.bc)
    yield dev
    dev.close()


def case_of(w, kinds) -> cases.Case:
    return cases.Case("bump462", w.mesh, w.scheme, None, kinds, w.bc)


def test_full_size_step_is_the_oracles(channel):
    w, u = channel
    m = w.mesh
    kinds = np.where(w.bc_kind == meshes.FARFIELD, float(meshes.STATE), w.bc_kind)
    meshes.set_boundary_kinds(m, kinds.astype(np.int32))
    try:
        print("full-size bump: device residual and step (state inflow)", flush=True)
        dev = DeviceStepper(m, w.scheme, "fp32+acc64", kinds, w.bc)
        r_dev = dev.residual(u)
        dev.upload(u)
        step = dev.step(cfl=w.cfl, local=True)
        u_dev = dev.download()
        dev.upload(u)
        cons_dev = dev.conservation()
        dev.close()
        print("full-size bump: oracle residual and step", flush=True)
        orc.set_threads(min(16, os.cpu_count() or 1))
        try:
            ora = Oracle(m, w.scheme, "fp32+acc64", kinds, w.bc)
            r_ora = ora.residual(u)
            np.testing.assert_array_equal(r_dev, r_ora)
            u_ora, norms, h, invalid = ora.step(u, cfl=w.cfl, local=True)
            cons_ora = ora.conservation(u)
        finally:
            orc.set_threads(1)
        assert step.invalid is None and invalid is None
        np.testing.assert_array_equal(u_dev, u_ora)
        np.testing.assert_array_equal(step.residuals, norms)
        assert step.time_step == h
        for key in ("E", "D", "boundary", "faces"):
            np.testing.assert_array_equal(cons_dev[key], cons_ora[key])
    finally:
        meshes.set_boundary_kinds(m, w.bc_kind.astype(np.int32))


def test_full_size_fp32_within_fp64_and_conservative(channel, stepper):
    w, u = channel
    print("full-size bump: fp32+acc64 and fp64 residuals (farfield)", flush=True)
    r32 = stepper.residual(u)
    stepper.upload(u)
    cons = stepper.conservation()
    dev64 = DeviceStepper(w.mesh, w.scheme, "fp64", w.bc_kind, w.bc)
    r64 = dev64.residual(u)
    dev64.close()
    units = cases.normwise_units(case_of(w, w.bc_kind), EPS32).reshape(-1)
    err = float((np.abs(r32 - r64) / units).max())
    assert err <= 500.0, err
    # the residual is not trivially small: the perturbation drives it far above the bound
    assert float((np.abs(r64) / units).max()) > 1e4
    cfo = w.mesh.cell_face_offsets
    mm = int(np.diff(cfo).max())
    assert np.all(np.abs(cons["E"]) <= mm * 2.0**-53 * cons["D"])


def test_full_size_runs_repeat_their_bits(channel, stepper):
    w, u = channel
    print("full-size bump: two runs of three steps", flush=True)
    runs = []
    for _ in range(2):
        stepper.upload(u)
        norms = [stepper.step(cfl=w.cfl, local=True).residuals for _ in range(3)]
        runs.append((np.array(norms), stepper.download()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_full_size_tetrahedra_residual_is_the_oracles():
    """The manufactured-solution tetrahedra at their largest level (128^3 x 6 = 12,582,912 jittered
    Kuhn tetrahedra, sources, the exact state on the boundary), second-order HLLC with the limiter
    (K = 5), fp32+acc64: the device residual of the manufactured state equals the oracle's bit for
    bit."""
    print("full-size tetrahedra 128^3 x 6: building the mesh", flush=True)
    w = workloads.tets(128, "fp32+acc64")
    dev = DeviceStepper(w.mesh, w.scheme, "fp32+acc64", w.bc_kind, w.bc, w.source)
    r_dev = dev.residual(w.u)
    dev.close()
    print("full-size tetrahedra: oracle residual", flush=True)
    orc.set_threads(min(16, os.cpu_count() or 1))
    try:
        r_ora = Oracle(w.mesh, w.scheme, "fp32+acc64", w.bc_kind, w.bc, w.source).residual(w.u)
    finally:
        orc.set_threads(1)
    np.testing.assert_array_equal(r_dev, r_ora)
    assert np.abs(r_dev).max() > 0.0


def test_full_size_sweep_batch_steps_are_each_cases_steps():
    """The 64-case sweep at its full size (seed 1: 6,496,211 cells of 20k-500k-cell channels) in
    one batch stepper: three SSP-RK3 steps on local time steps give the smallest, the largest and
    two transonic cases the norms and states they get stepped alone (fp32+acc64, bit for bit)."""
    from autan_amd import batch, sweep

    print("full-size sweep: building 64 channels", flush=True)
    specs = sweep.specs(64, seed=1)
    cs = [sweep.build(s, "fp32+acc64") for s in specs]
    sizes = [c.mesh.n_cells for c in cs]
    assert sum(sizes) == 6496211
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 10, 59})
    b = batch.BatchStepper(cs, "fp32+acc64")
    b.upload()
    alone = {j: DeviceStepper(cs[j].mesh, cs[j].scheme, "fp32+acc64", cs[j].bc_kind, cs[j].bc) for j in pick}
    for j, d in alone.items():
        d.upload(cs[j].u)
    for _ in range(3):
        r, h, inv = b.step(cfl=0.8, local=True)
        for j, d in alone.items():
            s = d.step(cfl=0.8, local=True)
            np.testing.assert_array_equal(r[j], s.residuals)
            assert inv[j] is None and s.invalid is None
    got = b.download()
    for j, d in alone.items():
        np.testing.assert_array_equal(got[j], d.download())
        d.close()
    b.close()
