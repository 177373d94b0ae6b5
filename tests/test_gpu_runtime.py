"""GPU parity of the runtime seam around the stepper: wall loads, residual scales and steady runs,
each through the C ABI against the oracle (or a closed form) on the same inputs."""

from __future__ import annotations

import math

import numpy as np
import pytest

import cases
from autan_amd import steady
from autan_amd.stepper import DeviceStepper
from oracle import oracle as orc
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

LOAD_CASES = {
    "poly-ausm2-warped": lambda: cases.fixture_case("poly", "ausm+up"),
    "bump-hybrid-ausm2": lambda: cases.bump_case(16, 12, 8, hybrid=True),
    "tets-hllc2": lambda: cases.tets_case(5),
}


def build(case, policy):
    dev = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    ora = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    return dev, ora


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name", list(LOAD_CASES))
def test_face_force_matches_oracle_bit_for_bit(name, policy):
    case = LOAD_CASES[name]()
    dev, ora = build(case, policy)
    dev.upload(case.u)
    faces = np.arange(case.mesh.n_interior, case.mesh.n_faces, dtype=np.int32)
    got = dev.face_force(faces)
    want = ora.face_force(case.u, faces)
    np.testing.assert_array_equal(got, want)
    # The same state after a step: the force is that of the stepped state.
    dev.step(cfl=0.5)
    u1 = dev.download()
    np.testing.assert_array_equal(dev.face_force(faces[::3]), ora.face_force(u1, faces[::3]))


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_loads_are_exact_sums_of_the_face_forces(policy):
    case = cases.bump_case(16, 12, 8, hybrid=True)
    dev, ora = build(case, policy)
    dev.upload(case.u)
    m = case.mesh
    wall = np.nonzero(m.face_zone == 2)[0].astype(np.int32) + m.n_interior  # the lower wall, y = bump
    p_inf = 1.0 / 1.4
    center = np.array([1.5, 0.0, 0.5])
    force, moment = dev.loads(wall, center, p_inf)
    ff = ora.face_force(case.u, wall)
    s = m.face_area_vector.reshape(3, -1)[:, wall]
    f = ff - p_inf * s
    arm = m.face_centroid.reshape(3, -1)[:, wall] - center[:, None]
    mom = np.array([arm[1] * f[2] - arm[2] * f[1], arm[2] * f[0] - arm[0] * f[2], arm[0] * f[1] - arm[1] * f[0]])
    assert list(force) == [orc.exact_sum(f[k]) for k in range(3)]
    assert list(moment) == [orc.exact_sum(mom[k]) for k in range(3)]
    # Uniform pressure p_inf + dp over the wall: dp times the wall's area vectors (projected area 3 x 1).
    uni = cases.conserved(np.ones(m.n_cells), np.full(m.n_cells, 0.5), 0 * np.ones(m.n_cells), 0 * np.ones(m.n_cells),
                          np.full(m.n_cells, p_inf + 0.1))
    dev.upload(uni.reshape(-1))
    force, _ = dev.loads(wall, center, p_inf)
    assert m.kept_boundaries == 0  # no boundary face joins a gradient: a uniform field has none
    tol = 1e-12 if policy == "fp64" else 1e-6
    np.testing.assert_allclose(force, [0.0, -0.1 * 3.0, 0.0], rtol=0, atol=tol)


def test_residual_scales_closed_form():
    """Periodic hexahedra of [-10, 10]^3: every cell's surface is 6 h^2, so Phi_k = f_k 6 h^2."""
    case = cases.vortex_case(6)
    dev = DeviceStepper(case.mesh, case.scheme, "fp64", case.bc_kind, case.bc)
    rho, speed, p, g = 1.0, 0.5, 1.0 / 1.4, 1.4
    got = dev.residual_scales(rho, speed, p)
    h = 20.0 / 6
    c = math.sqrt(g * p / rho)
    q = speed + c
    mass = rho * q * 6 * h * h
    want = [mass, mass * q, mass * q, mass * q, mass * (c * c / (g - 1) + 0.5 * speed * speed)]
    np.testing.assert_allclose(got, want, rtol=1e-14)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_steady_run_matches_the_oracle_run_bit_for_bit(policy):
    """The same steady run (first-order start, the fp32+acc64 level rule with its floors, local
    time steps at CFL 0.8) over the device stepper and over the oracle: every reported residual
    and every decision agree."""
    case = cases.tets_case(4)
    dev, ora = build(case, policy)
    dev.upload(case.u)
    ref = np.array(case.scheme.reference)
    speed = math.sqrt(ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2) / ref[0]
    p = 0.4 * (ref[4] - 0.5 * ref[0] * speed**2)
    scales = dev.residual_scales(ref[0], speed, p)
    kw = dict(max_iterations=240, residual_orders=8.0, level_rule=policy == "fp32+acc64", scales=scales)
    got = steady.solve(dev, **kw)
    target = cases.OracleTarget(ora, case.u)
    want = steady.solve_with(target, **kw)
    assert got.iterations == want.iterations and got.status == want.status
    assert got.first_order_end == want.first_order_end and got.first_order_steps == want.first_order_steps
    np.testing.assert_array_equal(got.history, want.history)
    np.testing.assert_array_equal(got.provisional, want.provisional)
    assert got.first_order_steps > 0
    np.testing.assert_array_equal(dev.download(), target.u)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name", ["tets-hllc2", "bump-hybrid-ausm2"])
def test_conservation_sums_match_oracle(name, policy):
    """check_conservation's exact sums on the device: bit for bit with the oracle where no face
    calls exp/log, and E within its round-off bound everywhere."""
    case = LOAD_CASES[name]()
    dev, ora = build(case, policy)
    dev.upload(case.u)
    got = dev.conservation()
    want = ora.conservation(case.u)
    m = int(np.diff(case.mesh.cell_face_offsets).max())
    assert np.all(np.abs(got["E"]) <= m * 2.0**-53 * got["D"])
    if case.uses_exp_log:
        np.testing.assert_allclose(got["D"], want["D"], rtol=1e-6)
    else:
        for key in ("E", "D", "boundary", "faces"):
            np.testing.assert_array_equal(got[key], want[key])


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_device_resident_calls_give_the_host_calls_bits(policy):
    """upload_device / download_device / residual_device on device arrays, ordered against a
    caller's stream with no host synchronisation in between, give the bits of upload / download /
    residual; the null stream orders them too."""
    from hipmem import DeviceArray, Stream

    case = cases.tets_case(5)
    host = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    host.upload(case.u)
    host.step(cfl=0.5)
    want_u = host.download()
    want_r = host.residual(want_u)

    dev = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    n5 = 5 * case.mesh.n_cells
    src = np.ascontiguousarray(case.u, dtype=np.float64).copy()
    u, out, res, out0 = DeviceArray(n5), DeviceArray(n5), DeviceArray(n5), DeviceArray(n5)
    s = Stream()
    try:
        # the state reaches u on the caller's stream; the upload must wait for it
        u.copy_from_async(src, s.handle)
        dev.upload_device(u.ptr, s.handle)
        dev.step(cfl=0.5)
        dev.download_device(out.ptr, s.handle)
        dev.residual_device(out.ptr, res.ptr, s.handle)
        s.synchronize()
        np.testing.assert_array_equal(out.numpy(), want_u)
        np.testing.assert_array_equal(res.numpy(), want_r)
        dev.download_device(out0.ptr)  # the null stream
        np.testing.assert_array_equal(out0.numpy(), want_u)
    finally:
        for a in (u, out, res, out0):
            a.free()
        s.destroy()
