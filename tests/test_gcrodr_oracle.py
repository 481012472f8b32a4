"""CPU pinning of the GCRO-DR oracle (oracle/dense.py gcrodr, restating
la/solver.py:248-492) by what the method's mathematics fixes:

  * with no recycled space and a restart long enough for one cycle it is
    FGMRES: the same residual history as the oracle's FGMRES;
  * the harmonic Ritz pencil (2.16) with no recycled block,
    Hbar^T Hbar z = theta H^T z, has the eigenvalues of
    H + h^2 H^-T e_m e_m^T (the textbook characterization of harmonic Ritz
    values, Paige, Parlett & van der Vorst 1995);
  * only span(Y) matters: Y and Y R (R invertible) give the same iterates;
  * a right-hand side in range(A M Y) is solved by the projection alone,
    with zero Arnoldi iterations;
  * on a sequence of slowly varying systems with a few small outlying
    eigenvalues, recycling converges to numpy's solutions in fewer total
    iterations than the same restarted method without it.
"""
import numpy as np

from oracle import dense as od


def _outlier_matrix(n, rng):
    """Nonsymmetric, eigenvalues clustered near 1 plus five at 0.01 .. 0.05."""
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([0.01 * np.arange(1, 6), 1.0 + 0.3 * rng.random(n - 5)])
    T = np.diag(lam) + np.triu(0.05 * rng.standard_normal((n, n)), 1)
    return Q @ T @ Q.T


def test_first_cycle_is_fgmres():
    rng = np.random.default_rng(0)
    n = 120
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x1, i1 = od.gcrodr(lambda v: A @ v, b, rtol=1e-10, restart=80)
    x2, i2 = od.fgmres(lambda v: A @ v, b, rtol=1e-10, restart=80)
    assert i1["iterations"] == i2["iterations"]
    h1, h2 = np.array(i1["history"]), np.array(i2["history"])
    assert np.allclose(h1, h2, rtol=1e-8, atol=1e-15)
    assert np.linalg.norm(x1 - x2) <= 1e-9 * np.linalg.norm(x2)


def test_harmonic_ritz_pencil_matches_textbook_form():
    rng = np.random.default_rng(1)
    m = 12
    Hbar = np.triu(rng.standard_normal((m + 1, m)), -1)
    W = np.zeros((m + 1, m))
    W[:m, :m] = np.eye(m)
    import scipy.linalg as sl
    theta = sl.eig(Hbar.T @ Hbar, Hbar.T @ W, right=False)
    H = Hbar[:m]
    h = Hbar[m, m - 1]
    em = np.zeros(m)
    em[-1] = 1.0
    ref = np.linalg.eigvals(H + h ** 2 * np.linalg.solve(H.T, em)[:, None] * em[None, :])
    # the same multiset: every value of one set within rounding of the other
    dist = np.abs(theta[:, None] - ref[None, :])
    scale = np.maximum(np.abs(ref), 1.0)[None, :]
    assert np.all((dist / scale).min(axis=1) < 1e-8) and np.all((dist / scale).min(axis=0) < 1e-8)
    # the span routine keeps the smallest-magnitude ones
    P = od.harmonic_ritz_span(Hbar, W, 3)
    assert P.shape[0] == m and 3 <= P.shape[1] <= 4
    assert np.allclose(P.T @ P, np.eye(P.shape[1]), atol=1e-12)


def test_only_the_span_of_y_matters():
    rng = np.random.default_rng(2)
    n = 150
    A = _outlier_matrix(n, rng)
    b = rng.standard_normal(n)
    _, info = od.gcrodr(lambda v: A @ v, b, rtol=1e-10, restart=25, dim=6)
    Y = info["Y"]
    Rm = rng.standard_normal((Y.shape[1], Y.shape[1])) + 3 * np.eye(Y.shape[1])
    b2 = b + 0.1 * rng.standard_normal(n)
    xa, ia = od.gcrodr(lambda v: A @ v, b2, Y=Y, rtol=1e-10, restart=25, dim=6)
    xb, ib = od.gcrodr(lambda v: A @ v, b2, Y=Y @ Rm, rtol=1e-10, restart=25, dim=6)
    assert ia["iterations"] == ib["iterations"]
    assert np.linalg.norm(xa - xb) <= 1e-8 * np.linalg.norm(xa)


def test_rhs_in_recycled_range_needs_no_iteration():
    rng = np.random.default_rng(3)
    n = 90
    A = np.eye(n) + 0.5 * rng.standard_normal((n, n)) / np.sqrt(n)
    Mi = np.diag(1.0 / (1.0 + rng.random(n)))
    Y = rng.standard_normal((n, 4))
    b = A @ (Mi @ (Y @ rng.standard_normal(4)))
    x, info = od.gcrodr(lambda v: A @ v, b, M=lambda v: Mi @ v, Y=Y, rtol=1e-10)
    assert info["iterations"] == 0 and info["setup_matvecs"] == 4
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-10 * np.linalg.norm(x)


def test_recycling_saves_iterations_on_a_sequence():
    rng = np.random.default_rng(4)
    n = 200
    A0 = _outlier_matrix(n, rng)
    E = 0.01 * rng.standard_normal((n, n)) / np.sqrt(n)
    Y = None
    with_rec, without = 0, 0
    for s in range(4):
        A = A0 + s * E
        b = rng.standard_normal(n)
        x, info = od.gcrodr(lambda v: A @ v, b, Y=Y, dim=8, rtol=1e-8, restart=20, maxiter=2000)
        Y = info["Y"]
        _, plain = od.gcrodr(lambda v: A @ v, b, Y=None, dim=0, rtol=1e-8, restart=20,
                             maxiter=2000)
        assert info["converged"]
        assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-6 * np.linalg.norm(x)
        with_rec += info["iterations"]
        without += plain["iterations"]
    print(f"iterations over 4 systems: recycled {with_rec}, restarted without {without}")
    assert with_rec < without
