"""The observed order of accuracy on the manufactured solution (the reference's own verification
gate: Roe or HLLC, unlimited, levels 12-32, docs/design/euler.md, Manufactured solution), through
the device stepper's steady runs."""

from __future__ import annotations

import pytest

from autan_amd import order

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", ["fp64", "fp32+acc64"])
def test_unlimited_hllc_is_second_order(policy):
    rep = order.study(levels=(12, 16, 24), policies=(policy,), flux="hllc", limiter=False, max_iterations=20000,
                      residual_orders=6.0 if policy == "fp64" else 3.0, log=lambda *_: None)
    for var, (fitted, pair) in rep["orders"][policy].items():
        assert 1.8 <= fitted <= 2.6, (var, fitted, pair)
        assert 1.7 <= pair <= 2.7, (var, fitted, pair)
