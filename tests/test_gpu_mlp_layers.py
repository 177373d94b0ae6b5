"""The warm-start network's kernels, each on its own inputs, to bounds derived from the arithmetic:

- every product on the matrix cores (BF16 operands, exact BF16 x BF16 products, FP32 accumulation of
  K terms) against its FP64 value from the device's own BF16 operands: |error| <= gamma_K sum |a b|
  with gamma_K = (K + 2) 2^-24 (the bound of any FP32 summation order);
- every epilogue that rounds to BF16 (the SiLU activations, the backward's dZ) against the BF16
  rounding of its exact value: equal, except where the exact value lies within the device's FP32
  error (the products' bound above and the hardware exponential and reciprocal, at most
  2e-6 + 2e-7 |z| relative, and for the SiLU derivative 2^-20 of the size of its terms, since it
  crosses zero) of a point halfway between two BF16 values, where either neighbour is right;
- the epilogues computed in FP32 without rounding to BF16 (input normalisation, the output scaling,
  the loss gradient) bit for bit against numpy's FP32 arithmetic;
- the weight and bias gradients (FP32 sums over the samples) against FP64 sums of the device's
  own dZ and activations, to gamma_n.
A transposed operand, a dropped bias, a wrong activation or derivative breaks these bounds by orders
of magnitude."""

from __future__ import annotations

import numpy as np
import pytest

from autan_amd.mlp import MLP
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(n, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1], np.cos(x[:, 2] * x[:, 3]), x[:, 4] ** 2, x[:, 5] - x[:, 6],
                  np.tanh(x[:, 0] + x[:, 6])], axis=1).astype(np.float32)
    return x, t


def trained(seed=0):
    """A network some Adam steps into training, so that its weights are not the initial ones."""
    m = MLP(7, 256, 4, 5, seed=seed)
    x, t = data(8192, 9)
    m.fit_normalization(x, t)
    m.train(x, t, epochs=2, batch=1024, lr=2e-3, seed=seed)
    return m


def bf16_between(lo: np.ndarray, hi: np.ndarray):
    """The BF16 roundings (of the FP32 roundings) of the interval ends: a value anywhere in
    [lo, hi] rounds to one of them or to a value between them (rounding is monotone)."""
    a, b = mo.bf16(np.minimum(lo, hi).astype(np.float32)), mo.bf16(np.maximum(lo, hi).astype(np.float32))
    return a, b


def assert_bf16_rounding(dev, exact, err, what, pinned):
    """dev (BF16 values) is the BF16 rounding of some value within err of exact; at least the
    fraction `pinned` of the values have one BF16 number in reach (the check is exact there)."""
    a, b = bf16_between(exact - err, exact + err)
    ok = (dev >= a) & (dev <= b)
    assert np.all(ok), (what, int((~ok).sum()), float(np.abs(dev - exact)[~ok].max()))
    assert np.mean(a == b) >= pinned, (what, float(np.mean(a == b)))


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def split(params, w):
    return mo.unpack(params, w)


@pytest.mark.parametrize("n", [200, 1024])
def test_forward_layers_within_their_rounding(n):
    m = trained(1)
    x, t = data(n, 3)
    w = m.widths()
    layers = split(m.params(), w)
    norm = [np.asarray(a, np.float32) for a in m.norm]
    # Layer 0's input: the normalised inputs in FP32, rounded to BF16, padded with zeros.
    h = m.trace(x, t, "input", 0).astype(np.float64)
    want = np.zeros_like(h)
    want[:, :7] = mo.bf16((x - norm[0]) / norm[1])
    np.testing.assert_array_equal(h, want)
    for l, (W, b) in enumerate(layers):
        wb = mo.bf16(W.astype(np.float32))
        z = m.trace(x, t, "z", l).astype(np.float64)
        exact = h @ wb.T + b
        bound = (w[l] + 2) * U32 * (np.abs(h) @ np.abs(wb).T + np.abs(b))
        assert np.all(np.abs(z - exact) <= bound), (l, float((np.abs(z - exact) / np.maximum(bound, 1e-300)).max()))
        if l < len(layers) - 1:
            nxt = m.trace(x, t, "input", l + 1).astype(np.float64)
            s = silu(z)
            assert_bf16_rounding(nxt, s, np.abs(s) * (2e-6 + 2e-7 * np.abs(z)) + 1e-30, f"silu of layer {l}", 0.99)
            h = nxt
    # The outputs: the last pre-activations scaled in FP32 (the fused inference gives the same bits).
    y = m.predict(x)
    zl = m.trace(x, t, "z", len(layers) - 1)[:, :5]
    np.testing.assert_array_equal(y, zl * norm[3] + norm[2])


@pytest.mark.parametrize("n", [200, 1024])
def test_backward_layers_within_their_rounding(n):
    m = trained(2)
    x, t = data(n, 5)
    w = m.widths()
    layers = split(m.params(), w)
    nl = len(layers)
    norm = [np.asarray(a, np.float32) for a in m.norm]
    loss, grads = m.gradients(x, t)
    # The loss gradient 2 (y - t) / (n d) in FP32, rounded to BF16 (zero in the padding).
    z = m.trace(x, t, "z", nl - 1)
    e = z[:, :5] - (t - norm[2]) / norm[3]
    g = np.float32(2.0) * e / np.float32(n * 5)
    dz = m.trace(x, t, "dz", nl - 1).astype(np.float64)
    want = np.zeros_like(dz)
    want[:, :5] = mo.bf16(g)
    np.testing.assert_array_equal(dz, want)
    assert loss == pytest.approx(float(np.mean(e.astype(np.float64) ** 2)), rel=1e-5)
    at = [0]
    for (W, b) in layers:
        at.append(at[-1] + W.size + b.size)
    for l in range(nl - 1, -1, -1):
        W, b = layers[l]
        wb = mo.bf16(W.astype(np.float32))
        h = m.trace(x, t, "input", l).astype(np.float64)
        # dW = dZ^T H and db = column sums of dZ, FP32 sums over the samples.
        gw = grads[at[l] : at[l] + W.size].reshape(W.shape).astype(np.float64)
        gb = grads[at[l] + W.size : at[l + 1]].astype(np.float64)
        bw = (n + 2) * U32 * (np.abs(dz).T @ np.abs(h))
        assert np.all(np.abs(gw - dz.T @ h) <= bw + 1e-30), l
        assert np.all(np.abs(gb - dz.sum(axis=0)) <= (n + 2) * U32 * np.abs(dz).sum(axis=0) + 1e-30), l
        if l > 0:
            # dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}), the product in FP32, the derivative with the
            # hardware exponential and reciprocal, rounded to BF16.
            zp = m.trace(x, t, "z", l - 1).astype(np.float64)
            c = dz @ wb
            cb = (w[l + 1] + 2) * U32 * (np.abs(dz) @ np.abs(wb))
            sg = silu_grad(zp)
            exact = c * sg
            # The derivative's FP32 evaluation s (1 + z (1 - s)) errs by a few units of its terms'
            # size, not of its value, which crosses zero near z = -1.28: an absolute part too.
            sl = 1.0 / (1.0 + np.exp(-zp))
            terms = sl + np.abs(zp) * sl * (1.0 - sl)
            err = cb * np.abs(sg) + np.abs(exact) * (4e-6 + 4e-7 * np.abs(zp)) + np.abs(c) * terms * 2.0 ** -20 + 1e-30
            nxt = m.trace(x, t, "dz", l - 1).astype(np.float64)
            # The product's own bound (not its value) reaches here, so fewer values are pinned.
            assert_bf16_rounding(nxt, exact, err, f"dz of layer {l - 1}", 0.5)
            dz = nxt
