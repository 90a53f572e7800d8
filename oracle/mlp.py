"""CPU reference of the learned tuner's cost model.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP network kernels
(dedispersion_amd/csrc/mlp.hip); the product never imports it.

The reference has no learned tuner (src/DedispersionTuning.cpp:201-427 times
every configuration), so there is no reference code to restate.  This module
states the network the HIP kernels implement -- the same flat parameter
layout, ReLU hidden layers, linear output and mean squared error -- in numpy
float64, and derives its gradient by hand (backpropagation).  The gradient is
pinned by central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np


def layout(inputs: int, hidden: int, layers: int) -> List[Tuple[int, int, int, int]]:
    """Per dense layer: (weight offset, fan_in, fan_out, bias offset)."""
    out, off = [], 0
    for l in range(layers + 1):
        a = inputs if l == 0 else hidden
        b = 1 if l == layers else hidden
        w = off
        off += a * b
        out.append((w, a, b, off))
        off += b
    return out


def param_count(inputs: int, hidden: int, layers: int) -> int:
    w, a, b, bo = layout(inputs, hidden, layers)[-1]
    return bo + b


def forward(params: np.ndarray, x: np.ndarray, inputs: int, hidden: int, layers: int):
    """Returns (y [n], activations [input, h1, ..., output])."""
    p = np.asarray(params, dtype=np.float64)
    h = np.asarray(x, dtype=np.float64)
    acts = [h]
    lay = layout(inputs, hidden, layers)
    for l, (w, a, b, bo) in enumerate(lay):
        z = h @ p[w:w + a * b].reshape(a, b) + p[bo:bo + b]
        h = np.maximum(z, 0.0) if l < layers else z
        acts.append(h)
    return h[:, 0], acts


def loss_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray, inputs: int, hidden: int, layers: int):
    """L = mean((y - t)^2) and dL/dparams in the flat layout."""
    p = np.asarray(params, dtype=np.float64)
    y, acts = forward(p, x, inputs, hidden, layers)
    n = len(t)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    grad = np.zeros_like(p)
    g = (2.0 * d / n)[:, None]  # dL/dz of the output layer
    lay = layout(inputs, hidden, layers)
    for l in range(layers, -1, -1):
        w, a, b, bo = lay[l]
        xin = acts[l]
        grad[w:w + a * b] = (xin.T @ g).ravel()
        grad[bo:bo + b] = g.sum(axis=0)
        if l > 0:
            g = (g @ p[w:w + a * b].reshape(a, b).T) * (acts[l] > 0)
    return loss, grad


def adam_step(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, 2015); returns new (params, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    c1, c2 = 1 - beta1 ** step, 1 - beta2 ** step
    return params - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
