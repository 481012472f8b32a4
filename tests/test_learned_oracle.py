"""CPU pinning of the learned preconditioner's oracle (oracle/learned.py).

The oracle's hand-written backward is checked against central finite
differences of its own forward + loss in float64, entry by entry on a random
subset of every parameter block; the kNN against a sort of all distances; the
Adam step against the update written out from Kingma & Ba (2015), Algorithm 1.
"""
import numpy as np
import pytest

from oracle import learned as ol


def _setup(P=37, k=16, seed=0):
    rng = np.random.default_rng(seed)
    theta = rng.normal(scale=0.3, size=ol.n_params())
    X = rng.normal(size=(P, ol.N_IN))
    pts = rng.normal(size=(P, 3))
    nbr = ol.knn(pts, k)
    t = rng.normal(size=(P, ol.N_OUT))
    return theta, X, nbr, t


def test_knn_matches_sorted_distances():
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(60, 3))
    nbr = ol.knn(pts, 8)
    for i in range(60):
        d = ((pts - pts[i]) ** 2).sum(1)
        d[i] = np.inf
        assert list(nbr[i]) == list(np.argsort(d, kind="stable")[:8])


@pytest.mark.parametrize("s", [1.0, 0.37])
def test_backward_matches_finite_differences(s):
    theta, X, nbr, t = _setup()
    y, cache = ol.forward(theta, X, nbr)
    g = ol.backward(theta, cache, y, t, s)
    rng = np.random.default_rng(2)
    o = 0
    h = 1e-6
    for name, shp in ol.SHAPES:
        size = int(np.prod(shp))
        for idx in rng.choice(size, size=min(size, 6), replace=False):
            q = o + idx
            tp, tm = theta.copy(), theta.copy()
            tp[q] += h
            tm[q] -= h
            fp = ol.loss(ol.forward(tp, X, nbr)[0], t, s)
            fm = ol.loss(ol.forward(tm, X, nbr)[0], t, s)
            fd = (fp - fm) / (2 * h)
            assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (name, idx, fd, g[q])
        o += size


def test_lowrank_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    P, B = 11, 3
    y = rng.normal(size=(P, ol.N_OUT))
    R = rng.normal(size=(B, 3 * P))
    T = rng.normal(size=(B, 3 * P))
    L, dy = ol.lowrank_loss(y, R, T)
    Phi, Psi = ol.bases(y)
    # the bases' layout: row 3p + a of Phi is y[p, aK:(a+1)K]
    K = ol.RANK
    assert np.array_equal(Phi[3 * 4 + 2], y[4, 2 * K:3 * K])
    assert np.array_equal(Psi[3 * 4 + 1], y[4, 3 * K + K:3 * K + 2 * K])
    h = 1e-6
    for idx in rng.choice(y.size, size=40, replace=False):
        yp, ym = y.copy(), y.copy()
        yp.flat[idx] += h
        ym.flat[idx] -= h
        fd = (ol.lowrank_loss(yp, R, T)[0] - ol.lowrank_loss(ym, R, T)[0]) / (2 * h)
        assert abs(fd - dy.flat[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8


def test_backward_catches_a_planted_error():
    """A dropped 1/P on the global-mean path is visible to the check above."""
    theta, X, nbr, t = _setup()
    y, cache = ol.forward(theta, X, nbr)
    g = ol.backward(theta, cache, y, t, 1.0)
    bad = dict(cache)
    bad["P"] = 1            # pretend the mean were a sum in the backward only
    g_bad = ol.backward(theta, bad, y, t, 1.0)
    assert np.abs(g_bad - g).max() > 1e-3 * np.abs(g).max()


def test_adam_step_matches_algorithm_1():
    rng = np.random.default_rng(3)
    w, g = rng.normal(size=50), rng.normal(size=50)
    m, v = np.zeros(50), np.zeros(50)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    w1, m1, v1 = ol.adam_step(w, g, m, v, 1, lr, b1, b2, eps)
    # Algorithm 1 with its epsilon on sqrt(v_hat) rescaled to the folded form
    mh = (1 - b1) * g / (1 - b1)
    vh = (1 - b2) * g * g / (1 - b2)
    ref = w - lr * mh / (np.sqrt(vh) + eps / np.sqrt(1 - b2))
    assert np.allclose(w1, ref, rtol=1e-12, atol=1e-14)


def test_subspace_loss_gradient_matches_finite_differences():
    """The learned coarse space's ridge objective (|V - Phi X|^2 + lam |X|^2)
    / |V|^2 with X its minimiser: the analytic gradient -2 (V - Phi X) X^T
    equals central differences, and the loss of a Phi whose span holds V is
    only the ridge term."""
    rng = np.random.default_rng(8)
    P, B = 13, 4
    y = rng.normal(size=(P, ol.N_OUT))
    V = rng.normal(size=(B, 3 * P))
    lam = 0.3
    L, dy = ol.subspace_loss(y, V, lam)
    assert 0.0 < L < 1.0
    h = 1e-6
    for idx in rng.choice(y.size, size=60, replace=False):
        yp, ym = y.copy(), y.copy()
        yp.flat[idx] += h
        ym.flat[idx] -= h
        fd = (ol.subspace_loss(yp, V, lam)[0] - ol.subspace_loss(ym, V, lam)[0]) / (2 * h)
        assert abs(fd - dy.flat[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9
    assert not np.any(dy[:, 3 * ol.RANK:])
    Phi, _ = ol.bases(y)
    Vin = (Phi @ rng.normal(size=(ol.RANK, B))).T
    L0, _ = ol.subspace_loss(y, Vin, 1e-12)
    assert L0 < 1e-9
