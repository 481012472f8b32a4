"""CPU check of the 1-norm inverse estimator behind the LU condition
estimate (mbem_amd/la/condest.py) against LAPACK's own gecon on the same
factors (the reference's cond_estimate is gecon's, la/lu.py:241-248):
random well- and ill-conditioned matrices of several sizes, n = 1 and 2
included; the estimate never exceeds the exact 1-norm condition number."""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg.lapack import dgecon

from mbem_amd.la.condest import onenorm_inverse


@pytest.mark.parametrize("n", [1, 2, 7, 60, 400])
@pytest.mark.parametrize("kind", ["gauss", "graded", "near_singular"])
def test_matches_lapack_gecon(n, kind):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    if kind == "graded":
        A = A @ np.diag(np.logspace(0, 6, n))
    elif kind == "near_singular" and n > 1:
        A[:, -1] = A[:, 0] + 1e-9 * rng.standard_normal(n)
    lu, piv = sl.lu_factor(A)
    anorm = float(np.abs(A).sum(axis=0).max())
    rcond, info = dgecon(lu, anorm, norm="1")
    assert info == 0
    est = onenorm_inverse(lambda v: sl.lu_solve((lu, piv), v),
                          lambda v: sl.lu_solve((lu, piv), v, trans=1), n)
    got = 1.0 / ((1.0 / est) / anorm)
    assert abs(got - 1.0 / rcond) <= 1e-9 / rcond
    assert got <= np.linalg.cond(A, 1) * (1 + 1e-8)
