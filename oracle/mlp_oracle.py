"""CPU oracle of the learned-boundary MLP (96 -> 32 -> 32 -> 9, ReLU) and its gradient.

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels (cmakepro_cuda_seissim_v2_amd
learned_abc.mlp), never a product path.

``forward(params, X, bf16=True)`` follows the GPU's arithmetic: inputs, weights and the hidden
activations are rounded to BF16 (round to nearest even) before each product, products are
summed in FP64 (the GPU sums in FP32 on MFMA), biases are added in FP32. ``backward`` is plain
reverse-mode differentiation of the mean squared error through the same rounded values, in
FP64. With ``bf16=False`` the network is plain FP64, which the central-difference test pins.
"""
from __future__ import annotations

import numpy as np

SHAPES = (("W1", (96, 32)), ("b1", (32,)), ("W2", (32, 32)), ("b2", (32,)), ("W3", (32, 9)), ("b3", (9,)))


def bf16(a) -> np.ndarray:
    """Round float32 values to BF16 (round to nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def forward(params, X, bf16_round: bool = True):
    """(Y, H1, H2, cache) for rows X [n][96]."""
    w = unpack(params)
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf16_round else (lambda a: np.asarray(a, np.float64))
    x = rnd(X)
    W1, W2, W3 = rnd(w["W1"]), rnd(w["W2"]), rnd(w["W3"])
    p1 = x @ W1 + w["b1"]
    h1 = rnd(np.maximum(p1, 0.0))
    p2 = h1 @ W2 + w["b2"]
    h2 = rnd(np.maximum(p2, 0.0))
    y = h2 @ W3 + w["b3"]
    return y, h1, h2, (x, W1, W2, W3)


def loss_and_grad(params, X, T, bf16_round: bool = True, h1=None, h2=None):
    """Mean squared error over all n x 9 outputs and its gradient in the packed layout.

    h1 / h2 may be given (the GPU's rounded activations) so that the backward pass is compared
    on identical forward values."""
    y, fh1, fh2, (x, W1, W2, W3) = forward(params, X, bf16_round)
    w = unpack(params)
    if h1 is not None:
        fh1 = np.asarray(h1, np.float64)
        fh2 = np.asarray(h2, np.float64)
        y = fh2 @ W3 + w["b3"]
    n = X.shape[0]
    r = y - np.asarray(T, np.float64)
    loss = float(np.sum(r * r) / (n * 9))
    dy = 2.0 * r / (n * 9)
    g = {"W3": fh2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ W3.T) * (fh2 > 0)
    g["W2"] = fh1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (fh1 > 0)
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    grad = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return loss, grad, y
