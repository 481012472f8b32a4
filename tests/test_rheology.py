"""Laplace-route host logic (no GPU): rheology, the s grid, and the g split of
the basis coefficients that the batched Laplace solve rests on.

Pins (each from a closed form the mathematics fixes, not from the code under
test): Maxwell's mu(s) and the single pole of g(s) (rheology.py:45-46 and the
reference's viscoelastic gate clause f), the bulk modulus kept, the Burgers
plateau, and the reduction of t_coeffs / u_coeffs to affine forms in g
(docs/viscoelastic.md:68-133; gate clause a).
"""
import math

import numpy as np
import pytest

from mbem_amd import laplace, rheology as rh
from mbem_amd.kernels import KERNEL_T, KERNEL_U
from mbem_amd.kernels import basis as kb

MU, LAM = 30.0, 30.0


def test_s_grid_workload_count_and_spacing():
    tau = 1e19 / 3e10
    s = laplace.s_grid(0.01 * tau, 2e7 * tau)
    assert len(s) == 32
    assert all(a > b for a, b in zip(s, s[1:]))
    hi = math.log10(1.0 / (0.01 * tau)) + 0.5
    lo = math.log10(1.0 / (2e7 * tau)) - 0.5
    assert s[0] == pytest.approx(10 ** hi, rel=1e-13)
    assert s[-1] == pytest.approx(10 ** lo, rel=1e-13)
    d = np.diff(np.log10(s))
    assert np.allclose(d, d[0], rtol=0, atol=1e-12)
    assert -d[0] <= 1.0 / 3.0 + 1e-12          # at least 3 points a decade
    t_min, t_max = laplace.window(s)
    assert t_min == pytest.approx(0.01 * tau, rel=1e-12)
    assert t_max == pytest.approx(2e7 * tau, rel=1e-12)
    with pytest.raises(ValueError):
        laplace.s_grid(2.0, 1.0)


def test_maxwell_closed_forms():
    m = rh.Maxwell(tau=2.5)
    K = LAM + 2 * MU / 3
    g_e = rh.g_of(MU, LAM)
    tau2 = m.tau * (3 * K + 4 * MU) / (3 * K)
    assert m.shear_ratio(math.inf) == 1.0
    for s in (1e-3, 0.3, 7.0, 1e4, 0.5 + 2.0j):
        mu, lam = rh.moduli(m, MU, LAM, s)
        assert abs(lam + 2 * mu / 3 - K) <= 1e-13 * K
        assert abs(mu - MU * s / (s + 1 / m.tau)) <= 1e-13 * MU
        # g(s) = 2 + (g_e - 2) s / (s + 1/tau2): one pole, at tau (3K + 4mu)/(3K)
        assert abs(rh.g_of(mu, lam) - (2 + (g_e - 2) * s / (s + 1 / tau2))) <= 1e-13


def test_g_of_is_one_over_one_minus_nu():
    for mu, lam in ((30.0, 30.0), (1.0, 7.0), (3.0, 0.5)):
        nu = lam / (2 * (lam + mu))
        assert rh.g_of(mu, lam) == pytest.approx(1 / (1 - nu), rel=1e-15)


def test_burgers_plateau_and_validation():
    wide = rh.Burgers(tau_M=1e7, tau_K=1e-7, kelvin=1.5)
    assert wide.shear_ratio(math.inf) == 1.0
    assert abs(wide.shear_ratio(1.0) - 1.5 / 2.5) < 1e-6
    with pytest.raises(ValueError):
        rh.Maxwell(tau=-1.0)
    with pytest.raises(ValueError):
        rh.Burgers(tau_M=1.0, tau_K=0.0, kelvin=1.0)
    assert rh.from_spec({"kind": "maxwell", "tau": 3.0}) == rh.Maxwell(3.0)


def test_laplace_states_order_and_moduli():
    m = rh.Maxwell(tau=1.0)
    st = rh.laplace_states({"crust": (MU, LAM)}, {"crust": m}, [0.1, 10.0, 1.0])
    assert st[0].label == "elastic" and math.isinf(st[0].s) and st[0].materials == {}
    assert [x.s for x in st[1:]] == [10.0, 1.0, 0.1]
    for x in st[1:]:
        mat = x.materials["crust"]
        assert (mat.mu, mat.lam) == rh.moduli(m, MU, LAM, x.s)
    with pytest.raises(NotImplementedError):
        rh.laplace_states({"crust": (MU, LAM)}, {"crust": m}, [1.0], faults={"f": None})


@pytest.mark.parametrize("nu", [0.0, 0.1, 0.25, 0.4, 0.49])
@pytest.mark.parametrize("mu", [30.0, 2.0])
def test_coefficient_split_is_affine_in_g(nu, mu):
    from mbem_amd.viscoelastic import split_coeffs
    lam = 2 * mu * nu / (1 - 2 * nu)
    g = 1 / (1 - nu)
    c = kb.t_coeffs(mu, lam)
    # the T basis satisfies N2 = N1 + 3 N3: fold the N2 coefficient away
    reduced = np.array([c[0] + c[1], 0.0, 3 * c[1] + c[2], c[3], c[4], c[5]])
    want = split_coeffs(KERNEL_T, 0) + g * split_coeffs(KERNEL_T, 1)
    assert np.abs(reduced - want).max() <= 1e-14 * max(1.0, np.abs(c).max())
    u = kb.u_coeffs(mu, lam)
    want_u = (split_coeffs(KERNEL_U, 0) + g * split_coeffs(KERNEL_U, 1)) / mu
    assert np.abs(u - want_u).max() <= 1e-15 * np.abs(want_u).max() * 10


def test_split_near_fluid_limit_stays_order_one():
    from mbem_amd.viscoelastic import split_coeffs
    K, mu = 50.0, 1e-6
    lam = K - 2 * mu / 3
    c = kb.t_coeffs(mu, lam)
    assert np.abs(c[:3]).max() > 1e6            # individually huge ...
    g = rh.g_of(mu, lam)
    want = split_coeffs(KERNEL_T, 0) + g * split_coeffs(KERNEL_T, 1)
    reduced = np.array([c[0] + c[1], 3 * c[1] + c[2]])
    # ... and what multiplies N1 and N3 is O(1): digits are lost only in the sum
    assert np.abs(reduced - want[[0, 2]]).max() / np.abs(want[[0, 2]]).max() < 1e-8


def test_require_one_eps():
    from mbem_amd.cases import box_fault_model
    model, eps, _ = box_fault_model(4, fault_cells=(2, 4))
    rh.require_one_eps(model, 0.7, {"crust"})
    bad = dict(eps)
    bad["top"] = 1.3 * bad["sides"]
    with pytest.raises(ValueError):
        rh.require_one_eps(model, bad, {"crust"})
    rh.require_one_eps(model, bad, set())            # nothing relaxes: no check
