"""float64 reference of the selector MLP's layers, loss and Adam step.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
dtc-spmm_asplos24_amd/csrc/selector.hip.  The layer definitions are the
textbook ones (y = act(x W + b), W row-major [in x out]); the gradients are
pinned in tests/test_selector_data.py by central finite differences, and
Adam follows Kingma & Ba (2015), Algorithm 1.
"""
from __future__ import annotations

import numpy as np


def linear_forward(x, w, b, relu):
    z = x.astype(np.float64) @ w.astype(np.float64)
    if b is not None:
        z = z + b.astype(np.float64)
    return np.maximum(z, 0.0) if relu else z


def linear_backward(x, w, y, dy, relu):
    """(dx, dw, db) given the layer output y (for the ReLU mask) and dy."""
    dz = dy.astype(np.float64) * ((y > 0) if relu else 1.0)
    dw = x.astype(np.float64).T @ dz
    db = dz.sum(axis=0)
    dx = dz @ w.astype(np.float64).T
    return dx, dw, db


def mse(pred, target, mask, inv_count):
    d = (pred.astype(np.float64) - target) * mask
    return inv_count * float((d * d).sum()), 2.0 * inv_count * d


def adam(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def mlp_forward(x, params):
    """3-layer MLP: relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3."""
    (w1, b1), (w2, b2), (w3, b3) = params
    h1 = linear_forward(x, w1, b1, True)
    h2 = linear_forward(h1, w2, b2, True)
    return h1, h2, linear_forward(h2, w3, b3, False)


def mlp_loss_and_grads(x, target, mask, params):
    (w1, b1), (w2, b2), (w3, b3) = params
    inv = 1.0 / max(1.0, float(mask.sum()))
    h1, h2, y = mlp_forward(x, params)
    loss, dy = mse(y, target, mask, inv)
    dh2, dw3, db3 = linear_backward(h2, w3, y, dy, False)
    dh1, dw2, db2 = linear_backward(h1, w2, h2, dh2, True)
    _, dw1, db1 = linear_backward(x, w1, h1, dh1, True)
    return loss, [(dw1, db1), (dw2, db2), (dw3, db3)]
