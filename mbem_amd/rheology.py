"""Linear viscoelastic rheology as operational shear moduli (rheology.py).

A region's ``(mu, lam)`` is its unrelaxed elasticity.  A rheology gives the
shear modulus at Laplace parameter ``s`` as ``mu(s) = mu * shear_ratio(s)``;
the bulk modulus ``K = lam + 2 mu / 3`` does not relax.  ``s = inf`` is the
elastic state and returns a ratio of exactly 1.

Only what the Laplace-state solve needs is here: ``Maxwell`` and ``Burgers``
(rheology.py:31-75), ``moduli`` (:479-488), ``g_of`` (:491-494),
``laplace_states`` (:497-524) and ``require_one_eps`` (:527-551).  Power-law
rheologies and creeping faults are not part of the dense hot path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Mapping


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


KINDS = {"maxwell": Maxwell, "burgers": Burgers}


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
