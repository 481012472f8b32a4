"""Linear viscoelastic rheology as operational shear moduli (rheology.py).

A region's ``(mu, lam)`` is its unrelaxed elasticity.  A rheology gives the
shear modulus at Laplace parameter ``s`` as ``mu(s) = mu * shear_ratio(s)``;
the bulk modulus ``K = lam + 2 mu / 3`` does not relax.  ``s = inf`` is the
elastic state and returns a ratio of exactly 1.

The linear bodies the Laplace-state solve takes: ``Maxwell``, ``Burgers``,
``Andrade`` and ``ExtendedBurgers`` (rheology.py:31-276), with ``tails``
(:419-431) and ``warn_window`` (:434-476) for the fit in time, ``moduli``
(:479-488), ``g_of`` (:491-494), ``laplace_states`` (:497-524) and
``require_one_eps`` (:527-551).  Creeping faults (fault drag) are not part of
the dense hot path.

Andrade and a power-law band have no fastest relaxation inside the window
(their spectrum runs on to short times as rate^-alpha), so each reports the
power law as ``tail`` for ``laplace.Prony(tails=...)``.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Mapping

import numpy as np


@dataclass(frozen=True)
class Maxwell:
    """Spring and dashpot in series: mu(s) = mu s / (s + 1/tau), tau = eta/mu
    (rheology.py:31-46)."""
    tau: float
    kind: str = field(default="maxwell", init=False)

    def __post_init__(self):
        if not (self.tau > 0.0 and math.isfinite(self.tau)):
            raise ValueError(f"Maxwell tau {self.tau} must be positive")

    def shear_ratio(self, s):
        return 1.0 / (1.0 + 1.0 / (self.tau * s))


@dataclass(frozen=True)
class Burgers:
    """Maxwell element in series with a Kelvin element (rheology.py:49-75):
    1/mu(s) = 1/mu + 1/(mu tau_M s) + 1/(mu_K (1 + tau_K s)), kelvin = mu_K/mu."""
    tau_M: float
    tau_K: float
    kelvin: float
    kind: str = field(default="burgers", init=False)

    def __post_init__(self):
        for name in ("tau_M", "tau_K", "kelvin"):
            v = getattr(self, name)
            if not (v > 0.0 and math.isfinite(v)):
                raise ValueError(f"Burgers {name} {v} must be positive")

    def shear_ratio(self, s):
        return 1.0 / (1.0 + 1.0 / (self.tau_M * s) + 1.0 / (self.kelvin * (1.0 + self.tau_K * s)))


def _positive(owner: str, name: str, v) -> None:
    if not (v > 0.0 and math.isfinite(v)):
        raise ValueError(f"{owner} {name} {v} must be positive")


@dataclass(frozen=True)
class Andrade:
    """Elastic + power-law transient + steady creep (rheology.py:77-131):

        J(t) mu = 1 + (t / tau_A)^alpha + t / tau_M
        mu / mu(s) = s J^(s) mu = 1 + Gamma(1 + alpha) (tau_A s)^-alpha + 1 / (tau_M s)

    (the Laplace transform of t^alpha is Gamma(1 + alpha) / s^(1 + alpha)).
    alpha = 1 is Maxwell at the summed rate and has no tail."""
    tau_M: float
    tau_A: float
    alpha: float
    kind: str = field(default="andrade", init=False)

    def __post_init__(self):
        _positive("Andrade", "tau_M", self.tau_M)
        _positive("Andrade", "tau_A", self.tau_A)
        if not 0.0 < self.alpha <= 1.0:
            raise ValueError(f"Andrade alpha {self.alpha} must be in (0, 1]: "
                             f"a creep that slows, at most linear in time")

    @property
    def tail(self):
        """The exponent of the spectrum's power law at short times."""
        return None if self.alpha == 1.0 else float(self.alpha)

    @property
    def tail_from(self):
        """Below tau_A (transient strain under the elastic one) the power law
        is the spectrum; a fit's window must reach down to it."""
        return None if self.alpha == 1.0 else float(self.tau_A)

    def shear_ratio(self, s):
        return 1.0 / (1.0 + math.gamma(1.0 + self.alpha) * (self.tau_A * s) ** (-self.alpha)
                      + 1.0 / (self.tau_M * s))


@dataclass(frozen=True)
class ExtendedBurgers:
    """A Maxwell element in series with Kelvin elements (tau_K[i], kelvin[i]
    = mu_K,i / mu), rheology.py:134-276:

        J(t) mu = 1 + sum_i (1 - exp(-t / tau_K,i)) / kelvin_i + t / tau_M
        mu / mu(s) = 1 + sum_i 1 / (kelvin_i (1 + tau_K,i s)) + 1 / (tau_M s)

    ``band`` cuts the elements from a power-law absorption band and records
    it as ``power_law = (alpha, tau_L, tau_H)``; inside the band the body
    creeps as t^alpha, so a window that does not reach below tau_L fits it
    with a tail that stops where the band does."""
    tau_M: float
    tau_K: tuple
    kelvin: tuple
    power_law: tuple = None
    kind: str = field(default="extended_burgers", init=False)

    def __post_init__(self):
        _positive("ExtendedBurgers", "tau_M", self.tau_M)
        for name in ("tau_K", "kelvin"):
            v = tuple(float(x) for x in np.atleast_1d(getattr(self, name)))
            if not v or any(not (x > 0.0 and math.isfinite(x)) for x in v):
                raise ValueError(f"ExtendedBurgers {name} {v}: one positive number an element")
            object.__setattr__(self, name, v)
        if len(self.tau_K) != len(self.kelvin):
            raise ValueError(f"ExtendedBurgers: {len(self.tau_K)} times and {len(self.kelvin)} "
                             f"springs; an element is one of each")
        if self.power_law is not None:
            alpha, tau_L, tau_H = (float(x) for x in self.power_law)
            if not (0.0 < alpha < 1.0 and 0.0 < tau_L < tau_H and math.isfinite(tau_H)):
                raise ValueError(f"ExtendedBurgers power_law {self.power_law}: (alpha, tau_L, "
                                 f"tau_H) with alpha in (0, 1) and 0 < tau_L < tau_H")
            if not (tau_L <= min(self.tau_K) and max(self.tau_K) <= tau_H):
                raise ValueError(f"ExtendedBurgers: elements from {min(self.tau_K):g} to "
                                 f"{max(self.tau_K):g} are not a band from {tau_L:g} to {tau_H:g}")
            object.__setattr__(self, "power_law", (alpha, tau_L, tau_H))

    @classmethod
    def band(cls, tau_M: float, strength: float, alpha: float, tau_L: float, tau_H: float,
             per_decade: int = None) -> "ExtendedBurgers":
        """Elements for a retardation spectrum with density proportional to
        tau^(alpha - 1) on [tau_L, tau_H], total transient compliance
        ``strength`` (over the elastic one).  The integral over log tau is
        cut into decade panels (the last one short), each integrated by
        ``per_decade``-point Gauss-Legendre: node tau_i gets compliance
        proportional to its weight times tau_i^alpha, and the compliances are
        normalised to sum to ``strength`` exactly."""
        from . import defaults

        n = defaults.EXTENDED_BURGERS_PER_DECADE if per_decade is None else int(per_decade)
        if not (strength > 0.0 and math.isfinite(strength)):
            raise ValueError(f"a band's strength {strength} must be positive")
        if not (0.0 < alpha < 1.0 and 0.0 < tau_L < tau_H and math.isfinite(tau_H) and n >= 1):
            raise ValueError(f"a band is alpha in (0, 1), 0 < tau_L < tau_H and at least one "
                             f"node a decade: got {alpha}, {tau_L}, {tau_H}, {n}")
        a, b = math.log(tau_L), math.log(tau_H)
        npanel = max(1, math.ceil((b - a) / math.log(10.0) - 1e-9))
        nodes, wts = np.polynomial.legendre.leggauss(n)
        edges = np.linspace(a, b, npanel + 1)
        mid, half = 0.5 * (edges[1:] + edges[:-1]), 0.5 * np.diff(edges)
        log_tau = (mid[:, None] + half[:, None] * nodes[None, :]).ravel()
        tau = np.exp(log_tau)
        share = (half[:, None] * wts[None, :]).ravel() * tau ** alpha
        compliance = strength * share / share.sum()
        return cls(tau_M=tau_M, tau_K=tuple(tau), kelvin=tuple(1.0 / compliance),
                   power_law=(alpha, tau_L, tau_H))

    @property
    def strength(self) -> float:
        """Total transient compliance over the elastic one."""
        return float(sum(1.0 / k for k in self.kelvin))

    @property
    def tail(self):
        """(alpha, 1/tau_L, 1/tau_H): the band's power law between those
        RATES; None when the elements are no band's."""
        if self.power_law is None:
            return None
        alpha, tau_L, tau_H = self.power_law
        return alpha, 1.0 / tau_L, 1.0 / tau_H

    @property
    def tail_from(self):
        """The tau_A of the Andrade body the band is between its ends: with
        D(tau) = alpha tau^(alpha-1) / (tau_H^alpha - tau_L^alpha), the
        transient strength * Int D(tau) (1 - exp(-t/tau)) dtau tends to
        strength Gamma(1 - alpha) t^alpha / (tau_H^alpha - tau_L^alpha) for
        tau_L << t << tau_H, which is (t / tau_A)^alpha."""
        if self.power_law is None:
            return None
        alpha, tau_L, tau_H = self.power_law
        return float(((tau_H ** alpha - tau_L ** alpha)
                      / (self.strength * math.gamma(1.0 - alpha))) ** (1.0 / alpha))

    @property
    def tail_to(self):
        """Where the band stops at its short end."""
        return None if self.power_law is None else self.power_law[1]

    def shear_ratio(self, s):
        inv = 1.0 + 1.0 / (self.tau_M * s)
        for tau, k in zip(self.tau_K, self.kelvin):
            inv = inv + 1.0 / (k * (1.0 + tau * s))
        return 1.0 / inv


KINDS = {"maxwell": Maxwell, "burgers": Burgers, "andrade": Andrade,
         "extended_burgers": ExtendedBurgers}


def from_spec(spec):
    """A rheology from itself or from its dict form {"kind": ..., ...}
    (rheology.py:408-416)."""
    if not isinstance(spec, dict):
        return spec
    d = dict(spec)
    kind = d.pop("kind")
    if kind not in KINDS:
        raise KeyError(f"unknown rheology kind {kind!r}; known: {sorted(KINDS)}")
    return KINDS[kind](**d)


def tails(rheologies) -> tuple:
    """The distinct power-law tails of ``rheologies`` (objects or their dict
    form), sorted as ``Prony`` takes them: an exponent for a law with no end,
    (alpha, fast rate, slow rate) for a band (rheology.py:419-431)."""
    found = {getattr(from_spec(r), "tail", None) for r in rheologies}
    found = [float(a) if np.ndim(a) == 0 else tuple(float(x) for x in a)
             for a in found if a is not None]
    return tuple(sorted(found, key=lambda a: a if isinstance(a, tuple) else (a, math.inf, 0.0)))


def warn_window(regions, s) -> list:
    """Warn, and return the messages, for every region whose power law stops
    being the spectrum (``tail_from``) below the window's short end, unless
    the window reaches below the band itself (``tail_to``) and no tail is
    fitted (rheology.py:434-476).  ``regions``: name -> rheology."""
    from . import laplace

    t_min = laplace.window(s)[0]
    said = []
    for name, rheo in regions.items():
        rheo = from_spec(rheo)
        reach = getattr(rheo, "tail_from", None)
        stop = getattr(rheo, "tail_to", None)
        if stop is not None and stop >= t_min:
            continue
        if reach is not None and reach < t_min * (1.0 - 1e-9):
            said.append(f"region {name!r}: its power-law transient reaches the elastic strain "
                        f"at tau_A = {reach:g}, below the window's short end ({t_min:g}); the "
                        f"fit's tail is the wrong shape there and a stress is off by percents. "
                        f"Lay the window out from t_min <= tau_A")
            warnings.warn(said[-1], stacklevel=2)
    return said


@dataclass(frozen=True)
class Material:
    """Moduli of one region in one state (config.py:57-65)."""
    mu: float
    lam: float


@dataclass(frozen=True)
class State:
    """One material state of a Laplace run (config.py:150-179): ``materials``
    names every region that differs from the base; ``s = inf`` is elastic."""
    label: str
    materials: Mapping[str, Material] = field(default_factory=dict)
    s: float | None = None
    drag: Mapping[str, float] = field(default_factory=dict)


def moduli(rheology, mu, lam, s):
    """(mu(s), lam(s)) with K = lam + 2 mu / 3 kept: lam(s) = K - 2 mu(s)/3,
    never through Poisson's ratio (rheology.py:479-488)."""
    K = lam + 2.0 * mu / 3.0
    mu_s = mu * rheology.shear_ratio(s)
    return mu_s, K - 2.0 * mu_s / 3.0


def g_of(mu, lam):
    """g = 1/(1 - nu) = 2 (lam + mu)/(lam + 2 mu), in [1, 2] (rheology.py:491-494)."""
    return 2.0 * (lam + mu) / (lam + 2.0 * mu)


def laplace_states(base, rheology, s, faults=None) -> tuple:
    """The elastic state, then one state per real s, largest s first
    (rheology.py:497-524).  ``base``: region -> unrelaxed (mu, lam);
    ``rheology``: relaxing region -> rheology."""
    if faults:
        raise NotImplementedError("creeping faults (fault drag) are not part of the dense path")
    out = [State("elastic", {}, s=math.inf, drag={})]
    for k, sv in enumerate(sorted((float(v) for v in s), reverse=True), 1):
        out.append(State(f"s{k:02d}",
                         {name: Material(*moduli(from_spec(r), *base[name], sv))
                          for name, r in rheology.items()},
                         s=sv, drag={}))
    return tuple(out)


def require_one_eps(model, eps, regions) -> None:
    """Refuse a relaxing region whose closed boundary has more than one eps
    (rheology.py:527-551): a per-element width breaks the incompressible
    closed-surface identity, and K/mu(s) amplifies the defect."""
    from .kernels import basis as kb

    for r in model.regions:
        if r.name not in regions or not r.patches:
            continue
        lo = min(float(kb.resolve_patch_eps(eps, p).min()) for p in r.patches)
        hi = max(float(kb.resolve_patch_eps(eps, p).max()) for p in r.patches)
        if hi - lo > 1e-12 * hi:
            raise ValueError(
                f"region {r.name!r} relaxes, so its closed boundary "
                f"({', '.join(p.name for p in r.patches)}) needs ONE scalar eps; it has eps "
                f"from {lo:.4g} to {hi:.4g}. Pass eps as a number, or as a dict giving those "
                f"patches the same number")
