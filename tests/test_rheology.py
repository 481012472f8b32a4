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


def test_andrade_modulus_is_the_transform_of_its_creep():
    """mu / mu(s) = s J^(s) with J(t) mu = 1 + (t/tau_A)^alpha + t/tau_M:
    the transform by quadrature (rheology.py:77-131), 1e-9; the elastic
    instant is exact and alpha = 1 is Maxwell at the summed rate."""
    from scipy import integrate
    a = rh.Andrade(tau_M=50.0, tau_A=2.0, alpha=1.0 / 3.0)
    for s in (0.01, 0.3, 7.0):
        # s * Int exp(-s t) J(t) dt, split at 1/s for the quadrature
        f = lambda t: math.exp(-s * t) * (1.0 + (t / a.tau_A) ** a.alpha + t / a.tau_M)  # noqa: E731
        val = s * sum(integrate.quad(f, lo, hi, epsabs=0, epsrel=1e-12, limit=200)[0]
                      for lo, hi in ((0.0, 1.0 / s), (1.0 / s, 60.0 / s)))
        assert 1.0 / a.shear_ratio(s) == pytest.approx(val, rel=1e-9)
    assert a.shear_ratio(math.inf) == 1.0
    assert a.tail == pytest.approx(1.0 / 3.0) and a.tail_from == 2.0
    m = rh.Andrade(tau_M=50.0, tau_A=2.0, alpha=1.0)
    assert m.tail is None and m.tail_from is None
    for s in (0.01, 1.0):
        assert m.shear_ratio(s) == pytest.approx(rh.Maxwell(1.0 / (1 / 50.0 + 1 / 2.0)).shear_ratio(s),
                                                 rel=1e-14)
    for bad in (dict(tau_M=-1.0, tau_A=1.0, alpha=0.3), dict(tau_M=1.0, tau_A=1.0, alpha=1.5)):
        with pytest.raises(ValueError):
            rh.Andrade(**bad)


def test_extended_burgers_one_element_and_the_band():
    """One element is Burgers; ``band`` sums to its strength exactly, its
    modulus and creep converge on the continuous band (density
    tau^(alpha - 1) on [tau_L, tau_H], by quadrature in log tau) as the
    reference's table says (4 nodes a decade: creep 2.5e-5, mu(s) 1.6e-7 of
    the strength; checked at 1e-4 and 1e-6), and its tail_from is the
    Andrade tau_A the band's creep follows between its ends."""
    from scipy import integrate
    eb = rh.ExtendedBurgers(tau_M=40.0, tau_K=3.0, kelvin=2.5)
    bg = rh.Burgers(tau_M=40.0, tau_K=3.0, kelvin=2.5)
    for s in (0.05, 1.0, 20.0):
        assert eb.shear_ratio(s) == pytest.approx(bg.shear_ratio(s), rel=1e-14)
    alpha, lo, hi, strength = 0.3, 1e-2, 1e3, 2.0
    b = rh.ExtendedBurgers.band(tau_M=1e5, strength=strength, alpha=alpha, tau_L=lo, tau_H=hi)
    assert b.strength == pytest.approx(strength, rel=1e-14)
    assert len(b.tau_K) == 4 * 5 and b.tail == (alpha, 1.0 / lo, 1.0 / hi) and b.tail_to == lo
    norm = hi ** alpha - lo ** alpha

    def cont(g):      # strength * Int D(tau) g(tau) dtau over log tau
        h = lambda u: strength * alpha * math.exp(alpha * u) / norm * g(math.exp(u))  # noqa: E731
        return integrate.quad(h, math.log(lo), math.log(hi), epsabs=0, epsrel=1e-12, limit=400)[0]

    for s in (1e-3, 0.1, 10.0, 1e3):
        inv = 1.0 + cont(lambda tau: 1.0 / (1.0 + tau * s)) + 1.0 / (1e5 * s)
        assert abs(1.0 / b.shear_ratio(s) - inv) < 1e-6 * strength
    tK, kv = np.array(b.tau_K), np.array(b.kelvin)
    for t in (0.05, 3.0, 200.0):
        Jd = np.sum((1.0 - np.exp(-t / tK)) / kv)
        assert abs(Jd - cont(lambda tau: 1.0 - math.exp(-t / tau))) < 1e-4 * strength
    # deep inside the band the creep is Andrade's with tau_A = tail_from,
    # less the part of the power law below tau_L (all relaxed by t >> tau_L:
    # strength tau_L^alpha / (tau_H^alpha - tau_L^alpha)); the part above
    # tau_H is (t/tau_H)^(1 - alpha) smaller still
    wlo, whi = 1e-8, 1e8
    wide = rh.ExtendedBurgers.band(tau_M=1e5, strength=1.0, alpha=alpha, tau_L=wlo, tau_H=whi,
                                   per_decade=6)
    t = 1.0
    J = np.sum((1.0 - np.exp(-t / np.array(wide.tau_K))) / np.array(wide.kelvin))
    below = wlo ** alpha / (whi ** alpha - wlo ** alpha)
    assert J == pytest.approx((t / wide.tail_from) ** alpha - below, rel=1e-4)
    with pytest.raises(ValueError):
        rh.ExtendedBurgers(tau_M=1.0, tau_K=(1.0, 2.0), kelvin=(1.0,))
    with pytest.raises(ValueError):
        rh.ExtendedBurgers(tau_M=1.0, tau_K=(1.0,), kelvin=(1.0,), power_law=(0.3, 2.0, 5.0))


def test_tails_and_dict_forms():
    """``tails`` (rheology.py:419-431): distinct, sorted, from objects or
    their dict form; none for bodies with a fastest relaxation."""
    import dataclasses
    a = rh.Andrade(tau_M=50.0, tau_A=2.0, alpha=1.0 / 3.0)
    b = rh.ExtendedBurgers.band(tau_M=1e4, strength=2.0, alpha=0.3, tau_L=1e-3, tau_H=1e2)
    got = rh.tails([a, dataclasses.asdict(b), {"kind": "maxwell", "tau": 3.0}, a,
                    rh.Burgers(1.0, 2.0, 3.0)])
    assert got == ((0.3, 1e3, 1e-2), 1.0 / 3.0)
    assert rh.from_spec(dataclasses.asdict(b)) == b
    assert rh.tails([rh.Maxwell(2.0)]) == ()


def test_warn_window():
    """A window whose short end is above tau_A is warned of; a band the
    window reaches below is not (rheology.py:434-476)."""
    s = laplace.s_grid(10.0, 1e4)                    # t_min = 10
    a = rh.Andrade(tau_M=1e5, tau_A=1.0, alpha=1.0 / 3.0)
    with pytest.warns(UserWarning):
        said = rh.warn_window({"mantle": a}, s)
    assert len(said) == 1 and "mantle" in said[0]
    ok = rh.Andrade(tau_M=1e5, tau_A=20.0, alpha=1.0 / 3.0)
    band_above = rh.ExtendedBurgers.band(tau_M=1e5, strength=1.0, alpha=0.3, tau_L=20.0, tau_H=1e4)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert rh.warn_window({"a": ok, "b": band_above}, s) == []

