"""CPU oracle of the learned preconditioner's network (test infrastructure
only: imported by tests, never by the product path).

The network has no counterpart in the reference; this file restates the
network the GPU kernels implement (mbem_amd/csrc/learned.hip) in numpy,
forward and backward written out by hand, so the GPU gradients can be
checked against it, and this backward is itself pinned by central finite
differences in tests/test_learned_oracle.py.

Parameters (one flat float vector, this order):
  W1 (128, 16)  b1 (128)   point layer on the 16 input features
  W2 (128, 128) b2 (128)   message layer, averaged over the k neighbours
  W3 (128, 256) b3 (128)   on [h1 | m]
  W4 (128, 256) b4 (128)   on [h3 | mean(h3)]
  W5 (6K, 128)  b5 (6K)    output: two 3P x K bases Phi, Psi (K = 16)

The preconditioner built from the output is z = BJ(r) + Phi (Psi^T r), with
Phi[3p + a, k] = y[p, a K + k] and Psi[3p + a, k] = y[p, 3K + a K + k]; the
training loss over a batch of normalised pairs (r_b, t_b) is
(1/B) sum_b |Phi Psi^T r_b - t_b|^2 / |t_b|^2 (``lowrank_loss``).
"""
from __future__ import annotations

import numpy as np

WIDTH = 128
N_IN = 16
RANK = 16
N_OUT = 6 * RANK
SHAPES = [("W1", (WIDTH, N_IN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, 2 * WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, 2 * WIDTH)),
          ("b4", (WIDTH,)), ("W5", (N_OUT, WIDTH)), ("b5", (N_OUT,))]


def n_params() -> int:
    return int(sum(np.prod(s) for _, s in SHAPES))


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = theta[o:o + k].reshape(shp)
        o += k
    return out


def relu(x):
    return np.maximum(x, 0.0)


def forward(theta, X, nbr):
    """(y (P, 3), cache) for features X (P, 16) and neighbours nbr (P, k)."""
    p = unpack(theta)
    P = X.shape[0]
    h1 = relu(X @ p["W1"].T + p["b1"])
    a = relu(h1 @ p["W2"].T + p["b2"])
    m = a[nbr].mean(axis=1)
    c1 = np.concatenate([h1, m], axis=1)
    h3 = relu(c1 @ p["W3"].T + p["b3"])
    g = h3.mean(axis=0)
    W4a, W4b = p["W4"][:, :WIDTH], p["W4"][:, WIDTH:]
    h4 = relu(h3 @ W4a.T + (W4b @ g + p["b4"]))
    y = h4 @ p["W5"].T + p["b5"]
    cache = dict(X=X, nbr=nbr, h1=h1, a=a, m=m, c1=c1, h3=h3, g=g, h4=h4, P=P)
    return y, cache


def loss(y, t, s):
    """|s y - t|^2 / |t|^2"""
    d = s * y - t
    return float((d * d).sum() / max((t * t).sum(), 1e-300))


def bases(y):
    """(Phi, Psi), each 3P x K, from the network output y (P x 6K)."""
    P = y.shape[0]
    K = y.shape[1] // 6
    return y[:, :3 * K].reshape(3 * P, K), y[:, 3 * K:].reshape(3 * P, K)


def lowrank_loss(y, R, T):
    """(loss, dy) of (1/B) sum_b |Phi Psi^T r_b - t_b|^2 / |t_b|^2; R, T: B x 3P."""
    Phi, Psi = bases(y)
    B = R.shape[0]
    C = R @ Psi                                  # B x K
    pred = C @ Phi.T                             # B x 3P
    tt = np.maximum((T * T).sum(axis=1), 1e-300)
    d = pred - T
    loss = float(((d * d).sum(axis=1) / tt).sum() / B)
    dpred = 2.0 * d / (tt[:, None] * B)
    dPhi = dpred.T @ C                           # 3P x K
    dC = dpred @ Phi                             # B x K
    dPsi = R.T @ dC                              # 3P x K
    P = y.shape[0]
    K = y.shape[1] // 6
    dy = np.concatenate([dPhi.reshape(P, 3 * K), dPsi.reshape(P, 3 * K)], axis=1)
    return loss, dy


def subspace_loss(y, V, lam):
    """(loss, dy) of the learned coarse space's ridge objective
        L = (|V^T - Phi X|_F^2 + lam |X|_F^2) / |V|_F^2,
        X = (Phi^T Phi + lam I)^-1 Phi^T V^T,
    the part of the target rows V (B x 3P) outside span(Phi).  X minimises
    the bracket, so dL/dPhi = -2 (V^T - Phi X) X^T / |V|^2 exactly (lam held
    fixed); dy carries it in the Phi columns, the Psi columns zero."""
    Phi, _ = bases(y)
    K = Phi.shape[1]
    X = np.linalg.solve(Phi.T @ Phi + lam * np.eye(K), Phi.T @ V.T)
    R = V.T - Phi @ X
    v2 = float((V * V).sum())
    loss = (float((R * R).sum()) + lam * float((X * X).sum())) / v2
    dy = np.zeros_like(y)
    dy[:, :3 * K] = (-2.0 * R @ X.T / v2).reshape(y.shape[0], 3 * K)
    return loss, dy


def backward(theta, cache, y, t, s, dy=None):
    """Gradient with respect to theta (flat, same order) of ``loss(y, t, s)``,
    or, given ``dy``, of any loss whose gradient in y is dy."""
    p = unpack(theta)
    X, nbr, h1, a, c1, h3, g, h4, P = (cache[k] for k in
                                       ("X", "nbr", "h1", "a", "c1", "h3", "g", "h4", "P"))
    k = nbr.shape[1]
    if dy is None:
        tt = max(float((t * t).sum()), 1e-300)
        dy = 2.0 * s * (s * y - t) / tt
    grads = {}
    grads["W5"] = dy.T @ h4
    grads["b5"] = dy.sum(axis=0)
    dh4 = (dy @ p["W5"]) * (h4 > 0)
    W4a, W4b = p["W4"][:, :WIDTH], p["W4"][:, WIDTH:]
    dc = dh4.sum(axis=0)
    grads["W4"] = np.concatenate([dh4.T @ h3, np.outer(dc, g)], axis=1)
    grads["b4"] = dc
    dg = W4b.T @ dc
    dh3 = (dh4 @ W4a + dg[None, :] / P) * (h3 > 0)
    grads["W3"] = dh3.T @ c1
    grads["b3"] = dh3.sum(axis=0)
    dc1 = dh3 @ p["W3"]
    dh1 = dc1[:, :WIDTH].copy()
    dm = dc1[:, WIDTH:]
    da = np.zeros_like(a)
    np.add.at(da, nbr.reshape(-1), np.repeat(dm, k, axis=0) / k)
    da *= (a > 0)
    grads["W2"] = da.T @ h1
    grads["b2"] = da.sum(axis=0)
    dh1 += da @ p["W2"]
    dh1 *= (h1 > 0)
    grads["W1"] = dh1.T @ X
    grads["b1"] = dh1.sum(axis=0)
    return np.concatenate([grads[name].reshape(-1) for name, _ in SHAPES])


def knn(pts, k):
    """(P, k) nearest neighbours, self excluded, ties by index (brute force)."""
    P = pts.shape[0]
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    d[np.arange(P), np.arange(P)] = np.inf
    order = np.lexsort((np.broadcast_to(np.arange(P), (P, P)), d), axis=1)
    return order[:, :k]


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (Kingma & Ba 2015) with bias correction folded into the rate."""
    lr_t = lr * np.sqrt(1 - b2 ** t) / (1 - b1 ** t)
    gt = g + wd * w
    m = b1 * m + (1 - b1) * gt
    v = b2 * v + (1 - b2) * gt * gt
    w = w - lr_t * m / (np.sqrt(v) + eps)
    return w, m, v
