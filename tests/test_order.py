"""The order-of-accuracy study's arithmetic: the volume-weighted L2 error and the fitted orders."""

from __future__ import annotations

import numpy as np
import pytest

from autan_amd import order


def test_errors_weight_by_volume():
    vol = np.array([1.0, 3.0])
    exact = np.zeros((5, 2))
    u = exact.copy()
    u[:, 0] = 2.0  # error 2 on a cell of volume 1, 0 on one of volume 3
    np.testing.assert_allclose(order.errors(vol, u, exact), np.full(5, np.sqrt(4.0 / 4.0)))
    u[2] = [1.0, -1.0]
    assert order.errors(vol, u, exact)[2] == pytest.approx(1.0)


@pytest.mark.parametrize("p", [1.0, 2.0, 2.5])
def test_fit_order_recovers_a_power_law(p):
    h = np.array([1 / 64, 1 / 96, 1 / 128])
    e = 3.0 * h**p
    fitted, pair = order.fit_order(h, e)
    assert fitted == pytest.approx(p, abs=1e-12)
    assert pair == pytest.approx(p, abs=1e-12)


def test_finest_pair_is_the_two_smallest_h():
    h = [0.3, 0.1, 0.2]
    e = [0.09, 0.01 * 1.5, 0.04]  # the finest pair (0.1, 0.2) has order log(0.04/0.015)/log 2
    _, pair = order.fit_order(h, e)
    assert pair == pytest.approx(np.log(0.04 / 0.015) / np.log(2.0))


def test_one_level_has_no_order():
    with pytest.raises(ValueError):
        order.fit_order([0.1], [1e-3])
