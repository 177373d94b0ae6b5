"""GPU parity: the HIP library through its C ABI against the CPU oracle on the same inputs.

The bar (tests/physics/test_euler_device.cpp:68-211 in the reference): per cell and equation
|R_dev - R_host| <= 500 eps f_k A_i, and bit for bit where both sides add in the same order and
use no exp/log (the farfield and inlet conditions call them; the device's differ from glibc's in
their last bits). The device gathers each cell's faces in the reference's store-gather order, so
it is compared bit for bit with the oracle's gather and within the bound with its scatter.
"""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import mesh as meshes
from autan_amd.stepper import DeviceStepper, probe_math
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

EPS = {"fp64": np.finfo(np.float64).eps, "fp32+acc64": np.finfo(np.float32).eps}


def build(case: cases.Case, policy: str):
    dev = DeviceStepper(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    ora = Oracle(case.mesh, case.scheme, policy, case.bc_kind, case.bc, case.source)
    return dev, ora


def within(case, policy, a, b, factor=500.0):
    units = cases.normwise_units(case, EPS[policy]).reshape(-1)
    err = np.abs(a - b) / units
    return float(err.max())


CASES = {
    "vortex-roe1": lambda: cases.vortex_case(8, "roe"),
    "vortex-hllc2": lambda: cases.vortex_case(8, "hllc", True, True),
    "tets-hllc2": lambda: cases.tets_case(5),
    "tets-roe2-unlimited": lambda: cases.tets_case(5, "roe", True, False),
    "tets-ausm1": lambda: cases.tets_case(5, "ausm+up", False, False),
    "hex-roe2": lambda: cases.fixture_case("hex", "roe"),
    "hybrid-hllc2": lambda: cases.fixture_case("hybrid", "hllc"),
    "poly-ausm2-warped": lambda: cases.fixture_case("poly", "ausm+up"),
    "poly-roe2-state": lambda: cases.fixture_case("poly", "roe", kinds_by_zone=[meshes.STATE]),
    "bump-ausm2": lambda: cases.bump_case(),
    "bump-hybrid-ausm2": lambda: cases.bump_case(16, 12, 8, hybrid=True),
    "bump-hybrid-hllc2": lambda: cases.bump_case(16, 12, 8, "hllc", hybrid=True),
}


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name", list(CASES))
def test_residual_matches_oracle(name, policy):
    case = CASES[name]()
    dev, ora = build(case, policy)
    r_dev = dev.residual(case.u)
    r_gather = ora.residual(case.u, gather=True)
    assert np.all(np.isfinite(r_dev))
    if case.uses_exp_log:
        assert within(case, policy, r_dev, r_gather) <= 500.0
    else:
        np.testing.assert_array_equal(r_dev, r_gather)
    r_scatter = ora.residual(case.u, gather=False)
    assert within(case, policy, r_dev, r_scatter) <= 500.0
    # The limiter must matter where it is on: the residual moves by more than the bound without it.


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("name", ["tets-hllc2", "poly-ausm2-warped", "vortex-hllc2"])
def test_stage_arrays_match_oracle(name, policy):
    case = CASES[name]()
    dev, ora = build(case, policy)
    dev.residual(case.u)
    ora.residual(case.u)
    a, b = dev.arrays(), ora.arrays()
    for key in ("prim", "prim_b", "grad", "phi"):
        if case.uses_exp_log and key != "prim":
            np.testing.assert_allclose(a[key], b[key], rtol=1e-5, atol=1e-6, err_msg=key)
        else:
            np.testing.assert_array_equal(a[key], b[key], err_msg=key)


def test_limiter_changes_the_residual():
    """As the reference requires (tests/physics/test_euler_device.cpp:162-175): dropping the
    limiter moves R by far more than the parity bound, so a skipped limiter cannot pass."""
    case = cases.tets_case(5, "hllc", True, True)
    case.u = case.u * (1.0 + 0.05 * np.sin(np.arange(case.u.size)))
    dev, _ = build(case, "fp64")
    r_lim = dev.residual(case.u)
    case2 = cases.tets_case(5, "hllc", True, False)
    dev2, _ = build(case2, "fp64")
    r_unl = dev2.residual(case.u)
    assert within(case, "fp64", r_lim, r_unl) > 1000 * 500


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("name", ["vortex-roe1", "tets-hllc2", "hybrid-hllc2"])
def test_step_matches_oracle(name, policy, local):
    case = CASES[name]()
    dev, ora = build(case, policy)
    dev.upload(case.u)
    r = dev.step(cfl=0.5, local=local)
    u_dev = dev.download()
    u_ora, norms, h, inv = ora.step(case.u, cfl=0.5, local=local)
    assert r.invalid is None and inv is None
    if case.uses_exp_log:
        np.testing.assert_allclose(u_dev, u_ora, rtol=1e-9 if policy == "fp64" else 1e-5)
        np.testing.assert_allclose(r.residuals, norms, rtol=1e-9 if policy == "fp64" else 1e-4)
    else:
        np.testing.assert_array_equal(u_dev, u_ora)
        np.testing.assert_array_equal(r.residuals, norms)
        assert r.time_step == h


@pytest.mark.parametrize("variant", [("face_stage", 1), ("grad_occupancy", 7), ("grad_occupancy", 8),
                                     ("face_occupancy", 5), ("face_occupancy", 8), ("face_block", 512),
                                     ("face_block", 256), ("grad_nontemporal", 0)])
@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("name", ["tets-hllc2", "poly-ausm2-warped", "bump-hybrid-hllc2"])
def test_kernel_variants_give_the_same_bits(name, local, variant):
    """The autotuner's kernel variants (the face kernel staging its block's cells in LDS, the
    gradient and face kernels built for more waves per SIMD, smaller face blocks, the gradient's
    stencil stream loaded non-temporally or not) compute the same
    residual, norms and state as the default kernels, bit for bit (fp32+acc64)."""
    case = CASES[name]()
    plain, _ = build(case, "fp32+acc64")
    other, _ = build(case, "fp32+acc64")
    if variant == ("face_occupancy", 5) and case.mesh.all_planar:
        with pytest.raises(RuntimeError, match="face_occupancy"):
            other.tune(*variant)
        return
    other.tune(*variant)
    np.testing.assert_array_equal(other.residual(case.u), plain.residual(case.u))
    for dev in (plain, other):
        dev.upload(case.u)
    for _ in range(2):
        a = plain.step(cfl=0.5, local=local)
        b = other.step(cfl=0.5, local=local)
        np.testing.assert_array_equal(a.residuals, b.residuals)
        assert a.time_step == b.time_step
    np.testing.assert_array_equal(plain.download(), other.download())


def test_face_stage_needs_fp32():
    case = CASES["tets-hllc2"]()
    with pytest.raises(RuntimeError, match="face_stage"):
        build(case, "fp64")[0].tune("face_stage", 1)


@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("name", ["tets-hllc2", "bump-hybrid-ausm2"])
def test_graph_replays_give_the_eager_bits(name, local):
    """Steps replayed from a recorded HIP graph, with eager work (downloads, an evaluation, a
    residual) between the replays, give the eager steps' bits at every step."""
    case = CASES[name]()
    eager, _ = build(case, "fp32+acc64")
    graphed, _ = build(case, "fp32+acc64")
    graphed.tune("graph", 1)
    for d in (eager, graphed):
        d.upload(case.u)
    for k in range(5):
        a = eager.step(cfl=0.5, local=local)
        b = graphed.step(cfl=0.5, local=local)
        np.testing.assert_array_equal(a.residuals, b.residuals)
        assert a.time_step == b.time_step
        np.testing.assert_array_equal(eager.download(), graphed.download())
        if k == 2:
            np.testing.assert_array_equal(eager.evaluate().residuals, graphed.evaluate().residuals)
            np.testing.assert_array_equal(eager.residual(case.u), graphed.residual(case.u))


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_fixed_dt_steps_and_evaluate(policy):
    case = CASES["tets-hllc2"]()
    dev, ora = build(case, policy)
    dev.upload(case.u)
    u = case.u
    for _ in range(3):
        r = dev.step(dt=1e-3)
        u, norms, h, inv = ora.step(u, dt=1e-3)
        np.testing.assert_array_equal(r.residuals, norms)
        assert r.time_step == 1e-3 and h == 1e-3
    np.testing.assert_array_equal(dev.download(), u)
    e = dev.evaluate()
    res = ora.residual(u)
    assert e.invalid is None
    exact = np.sqrt(np.array([np.sum(np.square(res.reshape(5, -1)[k])) for k in range(5)]) / case.mesh.n_cells)
    np.testing.assert_allclose(e.residuals, exact, rtol=1e-12)


def test_invalid_cell_is_the_least_index():
    case = CASES["tets-hllc2"]()
    dev, ora = build(case, "fp32+acc64")
    u = case.u.reshape(5, -1).copy()
    bad = [17, 5, 40]
    u[0, bad] = -1.0
    dev.upload(u.reshape(-1))
    e = dev.evaluate()
    assert e.invalid == min(bad) == ora.first_invalid(u)


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
@pytest.mark.parametrize("kind", ["fixed", "local"])
def test_step_that_breaks_the_state_reports_the_oracles_cell(kind, policy):
    """A step too large for the state (a fixed step far above the CFL limit, or local steps at a
    CFL number of 40) leaves cells with a negative density or pressure: the step's own check (inside
    the last stage's face kernel) reports the least such cell, as the oracle's step does."""
    case = CASES["tets-hllc2"]()
    dev, ora = build(case, policy)
    dev.upload(case.u)
    if kind == "fixed":
        r = dev.step(dt=2.0)
        _, _, _, inv = ora.step(case.u, dt=2.0)
    else:
        r = dev.step(cfl=40.0, local=True)
        _, _, _, inv = ora.step(case.u, cfl=40.0, local=True)
    assert inv is not None
    assert r.invalid == inv


def test_second_order_switch_matches_oracle():
    case = CASES["tets-hllc2"]()
    dev, ora = build(case, "fp64")
    dev.set_second_order(False)
    ora.set_second_order(False)
    np.testing.assert_array_equal(dev.residual(case.u), ora.residual(case.u))
    dev.set_second_order(True)
    ora.set_second_order(True)
    np.testing.assert_array_equal(dev.residual(case.u), ora.residual(case.u))


def test_spectral_radius_matches_oracle():
    case = CASES["hybrid-hllc2"]()
    dev, ora = build(case, "fp64")
    dev.residual(case.u)
    ora.residual(case.u)
    np.testing.assert_allclose(dev.spectral_radius(), ora.spectral_radius(), rtol=1e-12)


def test_device_math_is_correctly_rounded():
    """sqrt and division on the device are correctly rounded in FP64 and FP32 (the parity above
    relies on it); exp and log are within a few ulps of glibc's."""
    rng = np.random.default_rng(3)
    x = rng.uniform(1e-3, 1e3, 100_000)
    y = rng.uniform(1e-3, 1e3, 100_000)
    d = probe_math(0, x, y)
    e = probe_math(0, x / 1e3, y)["exp"]
    assert np.max(np.abs(e - np.exp(x / 1e3)) / np.spacing(np.exp(x / 1e3))) <= 2
    np.testing.assert_array_equal(d["sqrt"], np.sqrt(x))
    np.testing.assert_array_equal(d["div"], x / y)
    assert np.max(np.abs(d["log"] - np.log(x)) / np.spacing(np.abs(np.log(x)) + 1e-300)) <= 2
    f = probe_math(1, x, y)
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    np.testing.assert_array_equal(f["sqrt"], np.sqrt(xf).astype(np.float64))
    np.testing.assert_array_equal(f["div"], (xf / yf).astype(np.float64))
    lf = np.log(xf.astype(np.float64))
    assert np.max(np.abs(f["log"] - lf) / np.spacing(np.abs(lf).astype(np.float32)).astype(np.float64)) <= 2
