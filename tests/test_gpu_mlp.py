"""The warm-start network's BF16 matrix-core kernels against the CPU oracle: forward outputs and
gradients (the oracle rounds where the device rounds; the device accumulates in FP32, the oracle
in FP64), and training that lowers the loss."""

from __future__ import annotations

import numpy as np
import pytest

from autan_amd.mlp import MLP
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(n, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1], np.cos(x[:, 2] * x[:, 3]), x[:, 4] ** 2, x[:, 5] - x[:, 6],
                  np.tanh(x[:, 0] + x[:, 6])], axis=1).astype(np.float32)
    return x, t


def net(seed=0):
    m = MLP(7, 256, 4, 5, seed=seed)
    x, t = data(2048, 9)
    m.fit_normalization(x, t)
    return m


def test_forward_matches_the_oracle():
    m = net()
    x, t = data(333, 1)
    y = m.predict(x)
    p = m.params()
    norm = [np.asarray(a, np.float64) for a in m.norm]
    yn, _ = mo.forward(p, m.widths(), x, norm, bf=True)
    want = yn[:, :5] * norm[3] + norm[2]
    scale = np.abs(want).max()
    err = np.abs(y - want).max() / scale
    assert err <= 2e-2, err


@pytest.mark.parametrize("n", [64, 1000])
def test_gradients_match_the_oracle(n):
    m = net(3)
    x, t = data(n, 2)
    loss, g = m.gradients(x, t)
    p = m.params()
    norm = [np.asarray(a, np.float64) for a in m.norm]
    l_o, g_o = mo.loss_and_gradients(p, m.widths(), x, t, norm, 5, bf=True)
    assert abs(loss - l_o) <= 1e-2 * l_o
    assert np.linalg.norm(g - g_o) <= 3e-2 * np.linalg.norm(g_o)
    # layer by layer, so that a wrong transpose in one product shows
    at = 0
    w = m.widths()
    for l in range(len(w) - 1):
        k, nn = w[l], w[l + 1]
        seg = slice(at, at + nn * k)
        at += nn * k + nn
        assert np.linalg.norm(g[seg] - g_o[seg]) <= 5e-2 * np.linalg.norm(g_o[seg]) + 1e-12, l


def test_training_lowers_the_loss():
    m = net(5)
    x, t = data(32768, 4)
    l0, _ = m.gradients(x[:4096], t[:4096])
    loss, sec = m.train(x, t, epochs=20, batch=1024, lr=2e-3, seed=1)
    l1, _ = m.gradients(x[:4096], t[:4096])
    assert l1 < 0.05 * l0, (l0, l1)
    assert sec > 0


@pytest.mark.parametrize("hidden,n", [(256, 333), (256, 70000), (128, 1000), (64, 65)])
def test_fused_inference_gives_the_layer_by_layer_bits(hidden, n):
    """The one-kernel forward (every layer's activations kept in LDS) and the training path's
    layer-by-layer products give the same outputs, bit for bit, ragged tiles included."""
    m = MLP(7, hidden, 4, 5, seed=3)
    x, t = data(2048, 9)
    m.fit_normalization(x, t)
    m.train(x, t, epochs=2, batch=256, lr=1e-3, seed=0)  # weights away from their initial values
    xs, _ = data(n, 4)
    m.set_fused(False)
    layered = m.predict(xs)
    m.set_fused(True)
    # every inference tile (samples per workgroup) gives the same bits
    for rows in (64, 96, 128):
        m.set_infer_rows(rows)
        np.testing.assert_array_equal(m.predict(xs), layered, err_msg=f"{rows} samples per workgroup")
        assert m.time_forward(4096, 2) > 0.0
