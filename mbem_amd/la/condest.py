"""1-norm estimate of an inverse from solves (Hager 1984, Higham 1988
Algorithm 4.1), in the order LAPACK's dlacn2 takes it, so that
``anorm * onenorm_inverse(...)`` is the condition number LAPACK's gecon
reports for the same factors -- the reference's ``DenseSolveReport.
cond_estimate`` (la/lu.py:241-248, backends/dense.py:552-553).

The solves run on the device (the LU factors stay in HBM); this loop only
sees the n-vectors they return.
"""
from __future__ import annotations

import numpy as np


def _sign(x):
    return np.where(x >= 0.0, 1.0, -1.0)


def onenorm_inverse(solve, solve_t, n: int, itmax: int = 5) -> float:
    """Estimate of ||A^-1||_1 given ``solve(v) = A^-1 v`` and ``solve_t(v) =
    A^-T v`` on float64 host vectors.  A lower bound, exact in most cases."""
    if n <= 0:
        return 0.0
    x = solve(np.full(n, 1.0 / n))
    if n == 1:
        return float(abs(x[0]))
    est = float(np.abs(x).sum())
    isgn = _sign(x)
    x = solve_t(isgn.copy())
    j = int(np.argmax(np.abs(x)))
    it = 2
    while True:
        e = np.zeros(n)
        e[j] = 1.0
        x = solve(e)
        estold = est
        est = float(np.abs(x).sum())
        s = _sign(x)
        if np.array_equal(s, isgn) or est <= estold:
            break
        isgn = s
        x = solve_t(s.copy())
        jlast = j
        j = int(np.argmax(np.abs(x)))
        if x[jlast] != abs(x[j]) and it < itmax:
            it += 1
            continue
        break
    # the alternating test vector (dlacn2's final step)
    i = np.arange(n)
    alt = np.where(i % 2 == 0, 1.0, -1.0) * (1.0 + i / (n - 1))
    x = solve(alt)
    temp = 2.0 * float(np.abs(x).sum()) / (3.0 * n)
    return max(est, temp)
