"""CPU oracle of the warm-start network (test infrastructure only): the forward pass, the mean
squared error and its gradients of autan_amd/csrc/mlp/mlp.hip restated in numpy.

Two modes: `bf16=True` rounds where the device rounds (inputs, weights, activations and the
backward's dZ to BF16, round to nearest even), accumulating in FP64 (the device accumulates in
FP32 on the matrix cores, so the two agree to accumulation round-off); `bf16=False` computes in
FP64 throughout, and its gradient is pinned by central finite differences of its own loss.
Parameters are laid out as the device's: per layer the weights (outputs x inputs, the first
layer's inputs padded to 32 and the last layer's outputs to 16) then the biases.
"""

from __future__ import annotations

import numpy as np

IN_PAD, OUT_PAD = 32, 16


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def widths(inputs: int, hidden: int, layers: int, outputs: int) -> list[int]:
    return [IN_PAD] + [hidden] * layers + [OUT_PAD]


def unpack(params: np.ndarray, w: list[int]):
    out, at = [], 0
    for l in range(len(w) - 1):
        k, n = w[l], w[l + 1]
        W = params[at : at + n * k].reshape(n, k).astype(np.float64)
        at += n * k
        b = params[at : at + n].astype(np.float64)
        at += n
        out.append((W, b))
    assert at == params.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, w, x, norm, bf=True):
    """Returns the normalised outputs (n x OUT_PAD) and the layers' (input, pre-activation)."""
    in_mean, in_scale = norm[0], norm[1]
    d = x.shape[1]
    X = np.zeros((x.shape[0], IN_PAD))
    X[:, :d] = (np.asarray(x, np.float32) - np.asarray(in_mean, np.float32)) / np.asarray(in_scale, np.float32)
    rnd = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    H = rnd(X)
    cache = []
    layers = unpack(params, w)
    for l, (W, b) in enumerate(layers):
        Z = H @ rnd(W).T + b
        cache.append((H, Z))
        H = rnd(silu(Z)) if l < len(layers) - 1 else Z
    return H, cache


def loss_and_gradients(params, w, x, t, norm, outputs, bf=True):
    """MSE of the first `outputs` normalised outputs and its gradient in parameter order."""
    n = x.shape[0]
    Y, cache = forward(params, w, x, norm, bf)
    T = (np.asarray(t, np.float32) - np.asarray(norm[2], np.float32)) / np.asarray(norm[3], np.float32)
    E = np.zeros_like(Y)
    E[:, :outputs] = Y[:, :outputs] - T
    loss = float(np.sum(E * E) / (n * outputs))
    rnd = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    dZ = rnd(2.0 * E / (n * outputs))
    layers = unpack(params, w)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        H, Z = cache[l]
        dW = dZ.T @ H
        db = dZ.sum(axis=0)
        grads.append((dW, db))
        if l > 0:
            dZ = rnd((dZ @ rnd(W)) * silu_grad(cache[l - 1][1]))
    flat = []
    for dW, db in reversed(grads):
        flat.append(dW.reshape(-1))
        flat.append(db)
    return loss, np.concatenate(flat)
