"""The warm-start network's CPU oracle pinned by finite differences and BF16 rounding's definition,
and the library's exports (no device needed)."""

from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as mo


def small_net(seed=0, inputs=3, hidden=64, layers=2, outputs=2):
    rng = np.random.default_rng(seed)
    w = mo.widths(inputs, hidden, layers, outputs)
    parts = []
    for l in range(len(w) - 1):
        k, n = w[l], w[l + 1]
        W = rng.normal(0, 1 / np.sqrt(k), (n, k))
        if l == 0:
            W[:, inputs:] = 0
        if l == len(w) - 2:
            W[outputs:, :] = 0
        parts += [W.reshape(-1), rng.normal(0, 0.1, n) * (np.arange(n) < (outputs if l == len(w) - 2 else n))]
    return np.concatenate(parts), w


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1).
    assert mo.bf16(np.array([1 + 2**-8]))[0] == 1.0
    assert mo.bf16(np.array([1 + 3 * 2**-8]))[0] == 1 + 2**-6
    assert mo.bf16(np.array([1 + 2**-8 + 2**-20]))[0] == 1 + 2**-7
    x = np.random.default_rng(1).normal(size=1000).astype(np.float32)
    r = mo.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2**-8)
    assert np.all(mo.bf16(r) == r)


def test_fp64_gradient_matches_finite_differences():
    p, w = small_net()
    rng = np.random.default_rng(2)
    x = rng.normal(size=(40, 3))
    t = rng.normal(size=(40, 2))
    norm = [np.zeros(3), np.ones(3), np.zeros(2), np.ones(2)]
    loss, g = mo.loss_and_gradients(p, w, x, t, norm, 2, bf=False)
    # The padded parameters have no effect: their gradients are zero.
    layers = mo.unpack(p, w)
    idx = rng.choice(np.flatnonzero(p != 0), 60, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mo.loss_and_gradients(q, w, x, t, norm, 2, bf=False)
        q[i] -= 2 * h
        lm, _ = mo.loss_and_gradients(q, w, x, t, norm, 2, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert len(layers) == len(w) - 1


def test_bf16_mode_stays_near_fp64():
    p, w = small_net(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(64, 3))
    t = rng.normal(size=(64, 2))
    norm = [np.zeros(3), np.ones(3), np.zeros(2), np.ones(2)]
    l64, g64 = mo.loss_and_gradients(p, w, x, t, norm, 2, bf=False)
    l16, g16 = mo.loss_and_gradients(p, w, x, t, norm, 2, bf=True)
    assert abs(l16 - l64) <= 0.05 * l64
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)
