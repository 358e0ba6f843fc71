"""CPU restatement of the result-size predictor (test infrastructure: only
tests/ and bench.py's CPU leg use it; the product trains and runs the
network with its own HIP kernels, virginian_amd/csrc/gpu/mlp.hip).

The network has no counterpart in the reference, so there is no reference
file to follow.  Its definition is the textbook one:
    h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = sigmoid(h2 W3 + b3)
    loss = mean((y - t)^2), mean(|y - t|), or the pinball loss of quantile
    tau, mean(max(tau (t - y), (1 - tau) (y - t))) (Koenker & Bassett 1978)
with parameters packed flat as W1[nin][nh] b1[nh] W2[nh][nh] b2[nh] W3[nh] b3,
and optionally `skip` times the first input added to the output's
pre-activation (the network then learns a correction to an estimate given
as that input's logit; the skip has no parameter).

Pinned by: central finite differences in float64 against `gradient`
(tests/test_mlp.py), which a wrong sign, a dropped ReLU mask or a
transposed weight fails; and a closed form for nh = 1 with zero weights.
"""
from __future__ import annotations

import numpy as np


def layout(nin: int, nh: int) -> dict:
    o = {}
    k = 0
    for name, size in (("W1", nin * nh), ("b1", nh), ("W2", nh * nh), ("b2", nh), ("W3", nh), ("b3", 1)):
        o[name] = (k, k + size)
        k += size
    o["n"] = k
    return o


def unpack(p: np.ndarray, nin: int, nh: int) -> dict:
    L = layout(nin, nh)
    g = {k: p[v[0]:v[1]] for k, v in L.items() if k != "n"}
    return {"W1": g["W1"].reshape(nin, nh), "b1": g["b1"], "W2": g["W2"].reshape(nh, nh), "b2": g["b2"],
            "W3": g["W3"].reshape(nh, 1), "b3": g["b3"]}


def forward(p: np.ndarray, x: np.ndarray, nin: int, nh: int, skip: float = 0.0):
    w = unpack(p, nin, nh)
    z1 = x @ w["W1"] + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ w["W3"] + w["b3"] + skip * x[:, :1]
    y = 1.0 / (1.0 + np.exp(-z3[:, 0]))
    return y, (h1, h2)


def _loss_of(y: np.ndarray, t: np.ndarray, absolute: bool, tau: float | None) -> float:
    if tau is not None:
        d = y - t
        return float(np.mean(np.where(d > 0, (1.0 - tau) * d, -tau * d)))
    return float(np.mean(np.abs(y - t))) if absolute else float(np.mean((y - t) ** 2))


def loss(p: np.ndarray, x: np.ndarray, t: np.ndarray, nin: int, nh: int, skip: float = 0.0,
         absolute: bool = False, tau: float | None = None) -> float:
    y, _ = forward(p, x, nin, nh, skip)
    return _loss_of(y, t, absolute, tau)


def gradient(p: np.ndarray, x: np.ndarray, t: np.ndarray, nin: int, nh: int, skip: float = 0.0,
             absolute: bool = False, tau: float | None = None):
    """(loss, gradient packed like p) by backpropagation; the loss is the
    mean squared error, with absolute=True the mean absolute error
    (derivative sign(y - t), 0 where y = t), with tau the pinball loss
    (derivative 1 - tau above the target, -tau below it, 0 on it)."""
    w = unpack(p, nin, nh)
    y, (h1, h2) = forward(p, x, nin, nh, skip)
    B = x.shape[0]
    if tau is not None:
        dl = np.where(y > t, 1.0 - tau, np.where(y < t, -tau, 0.0))
    else:
        dl = np.sign(y - t) if absolute else 2.0 * (y - t)
    d3 = (dl / B * y * (1.0 - y))[:, None]                 # dL/dz3
    gW3 = h2.T @ d3
    gb3 = d3.sum(axis=0)
    d2 = (d3 @ w["W3"].T) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ w["W2"].T) * (h1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return _loss_of(y, t, absolute, tau), g
