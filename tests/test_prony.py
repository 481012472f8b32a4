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
  * the guards: secular fits have no relaxed value or memory, repeated
    points and malformed tails are refused;
  * tails (laplace.py:21-49, 86-117): an Andrade body's creep, in closed
    form J(t) mu = 1 + (t/tau_A)^alpha + t/tau_M, is fitted to ~1e-3 with its
    tail and ~8x worse without; an extended-Burgers band's creep (the exact
    sum over its elements) to 2e-5 with the band's tail against 2e-2
    without; a band the window reaches below is dropped and the fit is the
    plain one bit for bit; ``memory`` splits the fit into kept poles and the
    part gone at once, with the elastic instant exact and later times within
    the bound of what was dropped;
  * a stored run (laplace.py:315-367): ``write_run`` then ``run_stack`` /
    ``run_fit`` give back the states in fit order (whatever order they were
    written in), the tails of the run's rheology, and a MANIFEST of the
    files' hashes.
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
    for bad in ((1.2,), ((0.3, 1.0, 2.0),), (-0.1,)):
        with pytest.raises(ValueError):
            Prony(s, tails=bad)
    with pytest.raises(ValueError):
        Prony([1.0, 2.0], tails=(0.2, 0.3))           # no pole left beside two tails
    with pytest.raises(ValueError):
        Prony([1.0, 1.0, 2.0])
    tau, C, gone = Prony(s).memory()
    assert np.array_equal(C, Prony(s).coefficients()) and not gone.any()


def _creep_stack(rheo, s):
    rows = np.concatenate([[np.inf], s])
    return np.array([1.0 / rheo.shear_ratio(r) for r in rows])    # s J^(s) mu


def test_andrade_creep_with_and_without_its_tail():
    from mbem_amd import rheology as rh
    s = np.array(s_grid(1.0, 1e3))
    t = np.logspace(0, 3, 60)
    for tau_a in (0.3, 1.0, 3.0):
        a = rh.Andrade(tau_M=300.0, tau_A=tau_a, alpha=1.0 / 3.0)
        st = _creep_stack(a, s)
        J = 1.0 + (t / tau_a) ** a.alpha + t / a.tau_M
        plain = Prony(s, secular=True)
        tailed = Prony(s, secular=True, tails=rh.tails([a]))
        assert tailed.tails == (a.alpha,) and tailed.tau.size == plain.tau.size - 1
        e0 = np.abs(plain.weights(t) @ st - J).max() / J.max()
        e1 = np.abs(tailed.weights(t) @ st - J).max() / J.max()
        print(f"Andrade tau_A {tau_a}: creep fit error {e0:.1e} plain, {e1:.1e} with the tail")
        assert e1 < 2e-3 and e1 < e0 / 5.0
        # u(0+) is still the elastic state exactly
        assert abs(tailed.weights(0.0)[0] @ st - st[0]) < 1e-9


def test_band_tail_and_a_band_below_the_window():
    from mbem_amd import rheology as rh
    s = np.array(s_grid(1.0, 1e3))
    t = np.logspace(0, 3, 60)
    for lo, tailed_bar in ((1e-3, 1e-4), (3.0, None)):
        b = rh.ExtendedBurgers.band(tau_M=1e4, strength=2.0, alpha=0.3, tau_L=lo, tau_H=1e2)
        st = _creep_stack(b, s)
        J = (1.0 + np.sum((1.0 - np.exp(-t[:, None] / np.array(b.tau_K)[None, :]))
                          / np.array(b.kelvin)[None, :], axis=1) + t / b.tau_M)
        pr = Prony(s, secular=True, tails=(b.tail,))
        plain = Prony(s, secular=True)
        e = np.abs(pr.weights(t) @ st - J).max() / J.max()
        e0 = np.abs(plain.weights(t) @ st - J).max() / J.max()
        print(f"band from {lo}: tails {pr.tails}, error {e:.1e} (plain {e0:.1e})")
        if tailed_bar is None:
            # the window reaches below the band: no comb pole carries any of
            # it, the tail is dropped and the fit is the plain one
            assert pr.tails == () and np.array_equal(pr.coefficients(), plain.coefficients())
            assert e < 1e-3
        else:
            assert pr.tails == (b.tail,) and e < tailed_bar and e0 > 100 * e


def test_memory_with_a_tail():
    s = np.array(s_grid(1.0, 1e3))
    pr = Prony(s, tails=(1.0 / 3.0,))
    rng = np.random.default_rng(3)
    st = rng.standard_normal((s.size + 1, 4))
    tau, C, gone = pr.memory()
    m = pr.tau.size
    kept = tau.size - m
    # one decade of comb poles at three points a decade
    assert 2 <= kept <= 4 and np.all(1.0 / tau[m:] <= 10.0 * (1 + 1e-9) / pr.tau.min())
    # the elastic instant: C summed plus gone
    inst = C.sum(axis=0) + gone
    assert np.abs(inst @ st - pr.weights(0.0)[0] @ st).max() < 1e-12 * np.abs(st).max() * 1e3
    # later: the exponentials of memory equal the fit within what was dropped
    t = np.array([1.0, 3.0, 30.0])
    mem = np.hstack([np.ones((t.size, 1)), np.exp(-t[:, None] / tau[None, :])]) @ C @ st
    amp = pr.coefficients()[1 + m] @ st
    dropped = np.exp(-t[:, None] * pr._rate[None, kept:]) @ pr._w[0, kept:]
    diff = np.abs(mem - pr.weights(t) @ st)
    assert np.all(diff <= (dropped[:, None] * np.abs(amp)[None, :]) * (1 + 1e-6) + 1e-12)
    assert np.all(dropped < 1e-3)
    # without a tail, memory is the fit itself and nothing is gone
    tau0, C0, g0 = Prony(s).memory()
    assert np.array_equal(C0, Prony(s).coefficients()) and not g0.any()


def test_run_directory_round_trip(tmp_path):
    import hashlib
    import json
    from mbem_amd import laplace, rheology as rh
    s = laplace.s_grid(1.0, 1e2)
    body = rh.Andrade(tau_M=300.0, tau_A=1.0, alpha=1.0 / 3.0)
    states = rh.laplace_states({"crust": (30.0, 30.0)}, {"crust": body}, s)
    rng = np.random.default_rng(5)
    field = rng.standard_normal((7, 3))
    sols = [{"u:top": field * (1.0 / body.shear_ratio(st.s)), "t:base": field[:2] * k}
            for k, st in enumerate(states)]
    order = [3, 0, 5, 1, 2, 4] + list(range(6, len(states)))       # written out of order
    d = laplace.write_run(tmp_path / "run", [states[i] for i in order], [sols[i] for i in order],
                          {"crust": body}, reports=[{"iterations": i} for i in order])
    s_back, stack = laplace.run_stack(d)
    assert np.array_equal(s_back, np.array(s))
    assert stack["u:top"].shape == (len(states), 7, 3)
    for k in range(len(states)):
        assert np.array_equal(stack["u:top"][k], sols[k]["u:top"])
        assert np.array_equal(stack["t:base"][k], sols[k]["t:base"])
    pr = laplace.run_fit(d, secular=True)
    assert pr.tails == (body.alpha,)
    ref = Prony(s, secular=True, tails=(body.alpha,))
    assert np.array_equal(pr.coefficients(), ref.coefficients())
    resolved = json.loads((d / "resolved.json").read_text())
    assert resolved["spec"]["laplace"]["rheology"]["crust"]["kind"] == "andrade"
    assert any(r["s"] == "inf" for r in resolved["states"])
    for line in (d / "MANIFEST").read_text().splitlines():
        digest, size, name = line.split()
        b = (d / name).read_bytes()
        assert hashlib.sha256(b).hexdigest() == digest and len(b) == int(size)
    assert {ln.split()[2] for ln in (d / "MANIFEST").read_text().splitlines()} == \
        {p.name for p in d.iterdir() if p.name != "MANIFEST"}
    # guards: colliding slugs, no elastic state, not a Laplace run
    with pytest.raises(ValueError):
        laplace.write_run(tmp_path / "bad", [states[1], rh.State("s01!", s=1.0)] + [states[0]],
                          [sols[1], sols[1], sols[0]], {"crust": body})
    with pytest.raises(ValueError):
        laplace.write_run(tmp_path / "bad2", states[1:3], sols[1:3], {"crust": body})
    (tmp_path / "plain").mkdir()
    (tmp_path / "plain" / "resolved.json").write_text(json.dumps({"states": [{"label": "base"}],
                                                                  "spec": {}}))
    with pytest.raises(ValueError):
        laplace.run_stack(tmp_path / "plain")
