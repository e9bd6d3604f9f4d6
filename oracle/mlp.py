"""FP64 reference of the learned BFS policy's MLP -- TEST INFRASTRUCTURE ONLY.

Restates sygraph_amd/csrc/policy.hip in numpy float64: an 8 -> 32 -> 32 -> 6
ReLU MLP (parameter layout W1[32][8] b1[32] W2[32][32] b2[32] W3[6][32] b3[6]),
the mean softmax cross-entropy, its analytic gradient and one Adam step, plus
the eight level features of policy.hpp.  The analytic gradient is pinned by
central finite differences in tests/test_policy.py; the HIP kernels are
checked against this module in tests/test_policy_gpu.py.  There is no
reference counterpart (the reference has no learned policy).
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 8, 32, 6
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_OUT, N_H)),
          ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 1542


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def forward(p: np.ndarray, X: np.ndarray):
    w = unpack(np.asarray(p, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0.0)
    s = h2 @ w["W3"].T + w["b3"]
    return s, (X, h1, h2)


def loss_and_grad(p: np.ndarray, X: np.ndarray, y: np.ndarray):
    w = unpack(np.asarray(p, dtype=np.float64))
    s, (X, h1, h2) = forward(p, X)
    B = len(y)
    mx = s.max(axis=1, keepdims=True)
    e = np.exp(s - mx)
    z = e.sum(axis=1, keepdims=True)
    prob = e / z
    loss = float(np.mean(np.log(z[:, 0]) + mx[:, 0] - s[np.arange(B), y]))
    ds = prob.copy()
    ds[np.arange(B), y] -= 1.0
    ds /= B
    d2 = (ds @ w["W3"]) * (h2 > 0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    g = {"W3": ds.T @ h2, "b3": ds.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ X, "b1": d1.sum(0)}
    return loss, pack(g)


def loss_only(p, X, y) -> float:
    return loss_and_grad(p, X, y)[0]


def finite_difference(p, X, y, idx, h=1e-6) -> float:
    q = np.array(p, dtype=np.float64)
    q[idx] += h
    lp = loss_only(q, X, y)
    q[idx] -= 2 * h
    lm = loss_only(q, X, y)
    return (lp - lm) / (2 * h)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def features(frontier_vertices, frontier_edges, unexplored, level, max_degree, n, m) -> np.ndarray:
    """The eight level features of sg::policy::feature (policy.hpp), in FP64 (the device
    evaluates the same formulas in FP32)."""
    fv = np.asarray(frontier_vertices, dtype=np.float64)
    fe = np.asarray(frontier_edges, dtype=np.float64)
    f = np.stack([
        np.log2(1 + fv) / 24, np.log2(1 + fe) / 32, np.log2(1 + np.asarray(unexplored, dtype=np.float64)) / 32,
        np.log2(1 + np.asarray(level, dtype=np.float64)) / 12, np.log2(1 + fe / np.maximum(fv, 1)) / 16,
        np.log2(1 + np.asarray(max_degree, dtype=np.float64)) / 24,
        np.full_like(fv, np.log2(1 + m / max(n, 1)) / 8), np.full_like(fv, np.log2(max(n, 1)) / 32)], axis=-1)
    return f.astype(np.float32)
