"""TEST INFRASTRUCTURE ONLY -- CPU checker of the task cost model.

Only tests/ may import this module.  It restates, in float64 numpy, the
perceptron the native library trains on the GPU (7 -> 64 -> 64 -> 1, ReLU
hidden layers, linear output, mean squared error, Adam with beta 0.9 / 0.999
and eps 1e-8), with the same flat parameter layout:

    W1[7][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3

The model has no counterpart in the reference (whose driver hands tasks out
in order through NXTVAL, ccsd_t_fused_driver.hpp:381), so there is no
reference vector to match.  ``loss_grad`` is pinned by central finite
differences of ``loss`` (tests/test_cost_model.py), and ``adam_step`` is the
published update of Kingma & Ba (2015, Algorithm 1).
"""
from __future__ import annotations

import numpy as np

IN, H = 7, 64
OW1 = 0
OB1 = OW1 + IN * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H
NPARAM = OB3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(IN, H), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def forward(p, x):
    """Returns y[n] and the activations (a1, a2)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.maximum(x @ W1 + b1, 0.0)
    a2 = np.maximum(a1 @ W2 + b2, 0.0)
    return a2 @ w3 + b3, a1, a2


def loss(p, x, t):
    y, _, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(p, x, t):
    """Mean squared error and its gradient by back-propagation (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    y, a1, a2 = forward(p, x)
    r = y - t
    dy = 2.0 * r / n
    d2 = np.outer(dy, w3) * (a2 > 0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    g = np.zeros(NPARAM)
    g[OW1:OB1] = (x.T @ d1).ravel()
    g[OB1:OW2] = d1.sum(0)
    g[OW2:OB2] = (a1.T @ d2).ravel()
    g[OB2:OW3] = d2.sum(0)
    g[OW3:OB3] = a2.T @ dy
    g[OB3] = dy.sum()
    return float(np.mean(r * r)), g


def adam_step(p, g, m1, m2, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update at step t (1-based); returns (p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
