"""CPU oracle of the learned sizing network (numpy, float64).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the library's HIP forward,
gradient and Adam kernels.  The network has no counterpart in the reference;
this restates the mathematics the kernels implement:

  h1 = relu(W1 x + b1)        W1: 64 x 32
  h2 = relu(W2 h1 + b2)       W2: 64 x 64
  y  = W3 . h2 + b3           W3: 64
  loss = mean((y - t)^2)

with relu'(s) = 1 for s > 0, else 0.  Parameter vector order:
W1 (row-major), b1, W2 (row-major), b2, W3, b3.  The analytic gradient is
pinned by central finite differences (tests/test_sizer_oracle.py).
"""
from __future__ import annotations

import numpy as np

NF, NH = 32, 64
SIZES = [("W1", NH * NF), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH), ("b3", 1)]
NPARAMS = sum(s for _, s in SIZES)  # 6337


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, o = {}, 0
    for name, s in SIZES:
        out[name] = w[o:o + s]
        o += s
    out["W1"] = out["W1"].reshape(NH, NF)
    out["W2"] = out["W2"].reshape(NH, NH)
    out["b3"] = out["b3"][0]
    return out


def forward(w, x):
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    s1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(s1, 0.0)
    s2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(s2, 0.0)
    return h2 @ p["W3"] + p["b3"]


def loss(w, x, t):
    y = forward(w, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(w, x, t):
    """(loss, gradient) by back-propagation."""
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = len(x)
    s1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(s1, 0.0)
    s2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(s2, 0.0)
    y = h2 @ p["W3"] + p["b3"]
    dy = 2.0 * (y - t) / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, p["W3"]) * (s2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ p["W2"]) * (s1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return float(np.mean((y - t) ** 2)), g


def finite_difference(w, x, t, idx, h=1e-6):
    w = np.asarray(w, dtype=np.float64).copy()
    out = []
    for i in idx:
        wp = w.copy()
        wm = w.copy()
        wp[i] += h
        wm[i] -= h
        out.append((loss(wp, x, t) - loss(wm, x, t)) / (2 * h))
    return np.array(out)


def adam_step(w, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t is the 1-based step number); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
