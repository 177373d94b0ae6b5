"""The steady run (first-order start, local time steps, convergence rule) of the native library,
driven through its callbacks by prescribed residual sequences whose outcome the reference's rules
fix (physics/include/autan/physics/first_order_start.hpp:22-137; runtime/src/run.cpp:2360-2389;
runtime/include/autan/runtime/round_off.hpp), and by the oracle on a small mesh."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd import steady
from oracle.oracle import Oracle


class Sequence:
    """A stepper whose norms are prescribed: first(n) at first order, second(n) at second order, n
    the number of steps taken so far."""

    def __init__(self, first, second=None, invalid_at=None):
        self.first = first
        self.second = second or first
        self.order = True
        self.n = 0
        self.orders = []
        self.invalid_at = invalid_at

    def norms(self):
        return np.asarray((self.second if self.order else self.first)(self.n), dtype=float)

    def step(self, cfl):
        assert cfl == steady.STEADY_CFL
        r = self.norms()
        self.orders.append(self.order)
        self.n += 1
        return r, None

    def evaluate(self):
        if self.invalid_at is not None and self.n == self.invalid_at:
            return self.norms(), 7
        return self.norms(), None

    def set_second_order(self, on):
        self.order = on


def ones(v):
    return np.full(5, float(v))


def test_first_order_start_ends_on_a_thousandfold_drop():
    """drop = r / r(1) of the first-order residual; the start ends on the step whose drop is below
    1e-3: 10^-n reaches 1e-4 on the fifth step."""
    s = Sequence(lambda n: ones(10.0**-n), lambda n: ones(2.0**-n))
    r = steady.solve_with(s, max_iterations=6, start_order="first")
    assert r.first_order_end == "drop" and r.first_order_steps == 5
    assert s.orders == [False] * 5 + [True]
    assert list(r.provisional) == [True] * 5 + [False]
    # Provisional iterations report the second-order residual of the same state.
    np.testing.assert_array_equal(r.history[:5], [ones(2.0**-n) for n in range(5)])
    np.testing.assert_array_equal(r.history[5], ones(2.0**-5))


def test_first_order_start_ends_on_a_stall():
    """fp64: a drop that sets no new low (below 0.9 of the last) for 100 steps has stalled."""
    s = Sequence(lambda n: ones(1.0))
    r = steady.solve_with(s, max_iterations=120, start_order="first")
    assert r.first_order_end == "stall" and r.first_order_steps == 101
    assert r.status == steady.BUDGET


def test_first_order_start_ends_at_the_cap():
    """A drop that keeps setting lows (every 50 steps) without reaching 1e-3 runs 201 steps."""
    s = Sequence(lambda n: ones(0.85 ** (n // 50)))
    r = steady.solve_with(s, max_iterations=205, start_order="first")
    assert r.first_order_end == "cap" and r.first_order_steps == 201


def test_first_order_start_ends_on_the_level_rule():
    """fp32+acc64: once the median of the last 100 first-order residuals is within 10 floors
    (0.5 * 2^-23 * Phi_k) for every watched equation."""
    s = Sequence(lambda n: ones(1e-9))
    r = steady.solve_with(s, max_iterations=105, level_rule=True, start_order="first", scales=ones(1.0))
    assert r.first_order_end == "level" and r.first_order_steps == 100
    # Without floors nothing is watched: the level rule never ends the start.
    s = Sequence(lambda n: ones(1e-9))
    r = steady.solve_with(s, max_iterations=205, level_rule=True, start_order="first")
    assert r.first_order_end == "cap"


def test_converges_at_the_residual_orders_drop():
    """Second order from the start: 2^-n reaches 1e-8 of iteration 1's residual at n = 27."""
    s = Sequence(lambda n: ones(0.5**n))
    r = steady.solve_with(s, max_iterations=100, start_order="second")
    assert r.converged and r.iterations == 28 and r.first_order_steps == 0
    np.testing.assert_array_equal(r.reference, ones(1.0))
    assert r.plateau_mask == 0


def test_momentum_components_share_the_largest_reference():
    """A momentum component drops from the largest momentum reference: momentum-y starting at 1e-3
    and halving every second step converges at 1e-8 of 1, not of 1e-3 (n = 34 instead of 54)."""

    def seq(n):
        r = ones(0.5**n)
        r[2] = 1e-3 * 0.5 ** (n / 2)
        return r

    r = steady.solve_with(Sequence(seq), max_iterations=100, start_order="second")
    assert r.converged and r.iterations == 35


def test_an_identically_zero_equation_is_exempt():
    def seq(n):
        r = ones(0.5**n)
        r[3] = 0.0
        return r

    r = steady.solve_with(Sequence(seq), max_iterations=100, start_order="second")
    assert r.converged and r.iterations == 28


def test_a_given_reference_replaces_iteration_one():
    """A warm start drops from its source's reference: an absolute threshold."""
    s = Sequence(lambda n: ones(1e-3 * 0.5**n))
    r = steady.solve_with(s, max_iterations=100, start_order="second", reference=ones(1.0))
    # 1e-3 * 2^-n <= 1e-8 from n = 17
    assert r.converged and r.iterations == 18


def test_round_off_plateau_converges_with_floors():
    """fp32+acc64 with floors: a flat residual within 10 floors converges once its level has set
    no new low for 100 steps (the level exists from the 100th step)."""
    s = Sequence(lambda n: ones(1e-9))
    r = steady.solve_with(s, max_iterations=300, level_rule=True, start_order="second", scales=ones(1.0))
    assert r.converged and r.iterations == 200 and r.plateau_mask == 0b11111
    # Without floors the same run only spends its budget.
    r = steady.solve_with(Sequence(lambda n: ones(1e-9)), max_iterations=300, level_rule=True, start_order="second")
    assert r.status == steady.BUDGET and r.iterations == 300


def test_divergence_is_reported():
    r = steady.solve_with(Sequence(lambda n: ones(1.0), invalid_at=0), max_iterations=10)
    assert r.status == steady.DIVERGED and r.diverged_cell == 7 and r.iterations == 0

    def nan_at_3(n):
        return ones(np.nan if n == 3 else 1.0)

    r = steady.solve_with(Sequence(nan_at_3), max_iterations=10, start_order="second")
    assert r.status == steady.DIVERGED and r.iterations == 3


def test_bad_options_are_refused():
    with pytest.raises(RuntimeError, match="max_iterations"):
        steady.solve_with(Sequence(lambda n: ones(1.0)), max_iterations=0)
    with pytest.raises(ValueError):
        steady.solve_with(Sequence(lambda n: ones(1.0)), start_order="third")


def test_oracle_steady_bump_run():
    """The oracle driven by the steady run on a small bump channel (fp64, AUSM+-up, farfield and
    slip): the first-order start ends, the residual drops three orders and the run converges."""
    case = cases.bump_case(12, 4, 2)
    ora = Oracle(case.mesh, case.scheme, "fp64", case.bc_kind, case.bc, case.source)
    target = cases.OracleTarget(ora, case.u)
    r = steady.solve_with(target, max_iterations=3000, residual_orders=3.0)
    assert r.converged, (r.iterations, r.history[-1] / r.reference)
    assert r.first_order_end in ("drop", "stall", "cap") and r.first_order_steps > 0
    ref = r.reference.copy()
    ref[1:4] = ref[1:4].max()  # momentum components drop from the largest momentum reference
    assert np.all(r.history[-1] <= ref * 1e-3)
    assert not np.all(r.history[-2] <= ref * 1e-3)
