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


def test_refinement_logic_on_the_cpu():
    """backends/dense.refine (la/lu.py:198-239) with scipy's FP32 LU as the
    single-precision solve: a well-conditioned system stops once the
    predicted remainder rate / (1 - rate) |d| / |x| is below 1e-10, and the
    actual error then is below that too and within a factor 10 of the
    prediction (the contraction is about single precision times the
    condition number); a Hilbert matrix (cond 1e16)
    does not contract and is reported not done, after at most one correction
    past the rate bound."""
    import scipy.linalg as sla
    import torch

    from mbem_amd.backends.dense import refine
    rng = np.random.default_rng(3)
    n = 200
    A = rng.standard_normal((n, n)) + 10.0 * np.eye(n)
    b = rng.standard_normal(n)
    lu = sla.lu_factor(A.astype(np.float32))

    def solve32(v, out):
        out.copy_(torch.from_numpy(sla.lu_solve(lu, v.numpy().astype(np.float32)).astype(float)))

    At = torch.from_numpy(A)
    x, rep = refine(solve32, lambda v, out: out.copy_(At @ v), torch.from_numpy(b))
    err = np.linalg.norm(x.numpy() - np.linalg.solve(A, b)) / np.linalg.norm(np.linalg.solve(A, b))
    assert rep["done"] and rep["passes"] <= 3 and rep["rate"] < 1e-3
    assert err <= 1e-10 and err <= 10.0 * max(rep["left"], 1e-15)
    H = sla.hilbert(14)
    luh = sla.lu_factor(H.astype(np.float32))
    Ht = torch.from_numpy(H)

    def solveh(v, out):
        out.copy_(torch.from_numpy(sla.lu_solve(luh, v.numpy().astype(np.float32)).astype(float)))

    _, reph = refine(solveh, lambda v, out: out.copy_(Ht @ v), torch.from_numpy(np.ones(14)))
    assert not reph["done"] and reph["rate"] > 1e-2 and reph["passes"] <= 2
