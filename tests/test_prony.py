"""The Laplace fit (mbem_amd/laplace.py Prony, after laplace.py:119-312)
checked against what the mathematics fixes, not against a retyping:

  * a spectrum on the fit's own relaxation times is recovered exactly:
    coefficients to 1e-9, u(t) = a0 + sum a_j exp(-t/tau_j) to 1e-10;
  * the elastic row pins u(0+): weights(0) @ stack is the elastic state;
  * a Maxwell element (relaxation time between grid points) relaxes as
    mu exp(-t / tau_M) to 1e-3 of mu inside the window, and its creep under
    the secular fit, (1 + t/tau_M) / mu, has the steady rate 1/(mu tau_M)
    to rounding;
  * the periodic weights are the geometric series they stand for
    (checked against a long explicit sum of shifted steps), and the locked
    fault's velocity averages over a cycle to minus the elastic response
    (Savage's back-slip statement), by quadrature;
  * the guards: secular fits have no relaxed value or memory, tails and
    repeated points are refused.
"""
import numpy as np
import pytest
from scipy import integrate

from mbem_amd.laplace import Prony, s_grid


def _stack_from(pr, a0, a, c=0.0):
    rows = np.concatenate([[np.inf], pr.s])
    with np.errstate(divide="ignore"):
        st = a0 + (1.0 / (1.0 + 1.0 / (rows[:, None] * pr.tau[None, :]))) @ a + c / rows
    return st


def test_exact_recovery_of_a_representable_spectrum():
    s = np.array(s_grid(1.0, 1e3))
    pr = Prony(s)
    rng = np.random.default_rng(0)
    a0, a = 0.7, rng.standard_normal(pr.tau.size)
    st = _stack_from(pr, a0, a)
    coef = pr.coefficients() @ st
    assert abs(coef[0] - a0) < 1e-9 and np.abs(coef[1:] - a).max() < 1e-9 * np.abs(a).max()
    t = np.array([0.0, 0.3, 5.0, 77.0, 900.0])
    ref = a0 + np.exp(-t[:, None] / pr.tau[None, :]) @ a
    assert np.abs(pr.weights(t) @ st - ref).max() < 1e-10 * np.abs(ref).max()
    # u(0+) is the elastic state itself
    assert abs(pr.weights(0.0)[0] @ st - st[0]) < 1e-10 * np.abs(st).max()
    assert abs(pr.relaxed_weights() @ st - a0) < 1e-9


def test_maxwell_relaxation_and_creep():
    mu, tau_m = 30.0, 37.3                      # between grid points
    s = np.array(s_grid(1.0, 1e3))
    pr = Prony(s)
    rows = np.concatenate([[np.inf], s])
    relax = mu / (1.0 + 1.0 / (rows * tau_m))   # s G^(s) of mu exp(-t/tau_m)
    t = np.logspace(0, 3, 40)
    err = np.abs(pr.weights(t) @ relax - mu * np.exp(-t / tau_m)).max() / mu
    print(f"Maxwell relaxation fit error in the window: {err:.1e}")
    assert err < 1e-3
    prs = Prony(s, secular=True)
    with np.errstate(divide="ignore"):
        creep = (1.0 + 1.0 / (rows * tau_m)) / mu   # s J^(s) of (1 + t/tau_m)/mu
    assert abs(prs.creep_rate_weights() @ creep - 1.0 / (mu * tau_m)) < 1e-10 / (mu * tau_m)
    assert np.abs(prs.weights(t) @ creep - (1.0 + t / tau_m) / mu).max() < 1e-8 / mu


def test_cycle_weights_are_the_geometric_series():
    s = np.array(s_grid(1.0, 1e3))
    pr = Prony(s)
    rng = np.random.default_rng(1)
    st = _stack_from(pr, 0.2, rng.standard_normal(pr.tau.size))
    T = 50.0
    t = np.array([0.5, 10.0, 49.0])
    n = np.arange(0, 4000)
    explicit = np.array([((pr.weights(tt + n * T) - pr.weights(n * T)) @ st).sum() for tt in t])
    assert np.allclose(pr.cycle_weights(t, T) @ st, explicit, rtol=1e-9, atol=1e-12)
    # and its rate is the series of rates
    rate = np.array([(pr.rate_weights(tt + n * T) @ st).sum() for tt in t])
    assert np.allclose(pr.cycle_rate_weights(t, T) @ st, rate, rtol=1e-9, atol=1e-12)


def test_locked_fault_velocity_averages_to_minus_elastic():
    s = np.array(s_grid(1.0, 1e3))
    pr = Prony(s)
    T = 200.0
    f = lambda t: pr.locked_rate_weights(t, T)[0]          # noqa: E731
    mean = integrate.quad_vec(f, 0.0, T, epsabs=1e-13, epsrel=1e-11,
                              points=list(np.logspace(-3, np.log10(T), 12)[:-1]))[0] / T
    e0 = np.zeros(pr.s.size + 1)
    e0[0] = 1.0
    assert np.abs(mean + e0).max() < 1e-8


def test_guards():
    s = np.array(s_grid(1.0, 1e2))
    prs = Prony(s, secular=True)
    with pytest.raises(ValueError):
        prs.relaxed_weights()
    with pytest.raises(ValueError):
        prs.memory()
    with pytest.raises(ValueError):
        prs.cycle_weights([1.0], 10.0)
    with pytest.raises(ValueError):
        Prony(s).creep_rate_weights()
    with pytest.raises(NotImplementedError):
        Prony(s, tails=(0.3,))
    with pytest.raises(ValueError):
        Prony([1.0, 1.0, 2.0])
    tau, C, gone = Prony(s).memory()
    assert np.array_equal(C, Prony(s).coefficients()) and not gone.any()
