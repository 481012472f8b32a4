"""Real Laplace collocation points (laplace.py:61-85) and the fit from a
stack of state solutions back to time (laplace.py:119-312).

For step data the transformed solution s u^(s) is the elastic solution at the
moduli of s; a run solves one elastic state per point of ``s_grid``.  A
relaxing body's response is a constant plus decaying exponentials (plus a
term linear in time for a creeping one), so s u^(s) is fitted by

    u~(s) = a0 + sum_j a_j s / (s + 1/tau_j)  [+ c / s]
    u(t)  = a0 + sum_j a_j exp(-t / tau_j)    [+ c t]

with the tau_j fixed at the reciprocals of the collocation points: the fit
is linear and involves no solution values, so it is one small matrix, the
same for every unknown; ``weights(t) @ stack`` is the response at t.  The
elastic state (s = inf) is the stack's first row and pins u(0+).  Applying
the weights to a device-resident stack is one library GEMM
(``apply_weights``).

A body with no fastest relaxation (a power-law tail: Andrade, a band that
runs on below the window) gets, per tail, one more unknown in place of the
fastest pole: the amplitude of a fixed comb of poles beyond the grid, on the
grid's own geometric spacing, weighted as the power law (laplace.py:21-49,
86-117).  ``run_stack`` / ``run_fit`` read a stored Laplace run
(laplace.py:315-367); ``write_run`` stores one in that layout.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import math
import pathlib
import re

import numpy as np

from . import defaults


def s_grid(t_min: float, t_max: float, per_decade: int = None, pad: float = None) -> tuple:
    """s = 1/t, log-spaced at ``per_decade`` points a decade over [t_min, t_max]
    widened by ``pad`` decades each side, largest first (laplace.py:61-76)."""
    per_decade = defaults.LAPLACE_POINTS_PER_DECADE if per_decade is None else per_decade
    pad = defaults.LAPLACE_PAD_DECADES if pad is None else pad
    if not (0.0 < t_min < t_max):
        raise ValueError(f"need 0 < t_min < t_max, got {t_min}, {t_max}")
    lo = math.log10(1.0 / t_max) - pad
    hi = math.log10(1.0 / t_min) + pad
    n = int(math.ceil((hi - lo) * per_decade)) + 1
    return tuple(float(v) for v in np.logspace(hi, lo, n))


def window(s, pad: float = None) -> tuple:
    """(t_min, t_max) that ``s_grid`` laid ``s`` out for (laplace.py:79-85)."""
    pad = defaults.LAPLACE_PAD_DECADES if pad is None else pad
    return 10.0 ** pad / max(s), 10.0 ** -pad / min(s)


def _as_tail(a) -> tuple:
    """(alpha, fastest rate, slowest rate) of a tail given as an exponent
    (no end: rates up to inf, down to 0) or with its band (laplace.py:86-94)."""
    if np.ndim(a) == 0:
        a = (float(a), math.inf, 0.0)
    else:
        a = tuple(float(x) for x in a)
        a = a + (math.inf, 0.0)[len(a) - 1:]
    if len(a) != 3 or not (0.0 < a[0] < 1.0 and 0.0 <= a[2] < a[1]):
        raise ValueError(f"a tail is an exponent in (0, 1), or that and a fastest and a "
                         f"slowest rate: {a}")
    return a


def _comb(tail, tau0: float, step: float, k: np.ndarray) -> tuple:
    """(w, rest) for the poles at rates step^k / tau0, k = 1..K (laplace.py:97-117).

    A spectrum rate^-alpha per unit log-rate puts, into the bin of log-width
    log(step) centred on pole k, a mass proportional to the measure of
    tau^alpha over the bin: tau^alpha evaluated at the bin's edges
    tau0 step^(-k +- 1/2).  Normalised by the bin of pole 0, an endless law
    gives w_k = step^(-alpha k) and the rest past pole K is the geometric
    remainder w_K / (step^alpha - 1).  A band keeps only the part of each bin
    between its ends (in tau: 1/fast .. 1/slow)."""
    alpha, fast, slow = tail
    if math.isinf(fast) and slow == 0.0:
        w = step ** (-alpha * k)
        return w, float(w[-1] / (step ** alpha - 1.0))
    tau_lo = 1.0 / fast
    tau_hi = 1.0 / slow if slow > 0.0 else math.inf
    upper = tau0 * step ** (0.5 - k)            # bin edges in tau, k-th pole
    lower = upper / step
    norm = tau0 ** alpha * (step ** (0.5 * alpha) - step ** (-0.5 * alpha))
    w = np.maximum(np.minimum(upper, tau_hi) ** alpha - np.maximum(lower, tau_lo) ** alpha,
                   0.0) / norm
    rest = max(min(lower[-1], tau_hi) ** alpha - tau_lo ** alpha, 0.0) / norm
    return w, float(rest)


class Prony:
    """The fit for one set of finite collocation points ``s``; a stack is
    ordered [elastic, s[0], s[1], ...] (laplace.py:119-312).

    ``secular`` adds the creep term c t and drops the slowest pole (the two
    are all but collinear over the window).  Each power-law tail (an exponent,
    or (alpha, fast rate, slow rate) for a band; ``rheology.tails``) replaces
    the fastest remaining pole by one unknown: the amplitude of the comb of
    poles beyond the grid, ``defaults.LAPLACE_TAIL_DECADES`` of them, shaped
    as the power law, plus the mass past the comb, which relaxes at once.  A
    band that stops short of every comb pole is no tail of this fit and is
    dropped (the grid then keeps that pole)."""

    def __init__(self, s, secular: bool = False, rcond: float = None, tails=()):
        s = np.asarray(s, float)
        if s.ndim != 1 or s.size < 2 or np.any(s <= 0) or np.any(~np.isfinite(s)):
            raise ValueError("s must be at least two finite positive values")
        if np.unique(s).size != s.size:
            raise ValueError("collocation points must be distinct")
        self.s = s
        self.secular = bool(secular)
        tau = np.sort(1.0 / s)
        if self.secular:
            tau = tau[:-1]
        step = (s.max() / s.min()) ** (1.0 / (s.size - 1))
        k = np.arange(1, int(math.ceil(defaults.LAPLACE_TAIL_DECADES / math.log10(step))) + 1)
        live = sorted({_as_tail(a) for a in tails})
        while True:
            if tau.size - len(live) < 1:
                raise ValueError(f"{s.size} points leave no pole beside {len(live)} tail(s)")
            self.tau = tau[len(live):]
            self._rate = step ** k / self.tau.min()
            combs = [_comb(a, self.tau.min(), step, k) for a in live]
            alive = [bool(w.any() or rest) for w, rest in combs]
            if all(alive):
                break
            live = [a for a, ok in zip(live, alive) if ok]
        self.tails = tuple(a[0] if a[1:] == (math.inf, 0.0) else a for a in live)
        self._w = np.array([w for w, _ in combs]).reshape(len(live), k.size)
        self._rest = np.array([r for _, r in combs])
        rows = np.concatenate([[np.inf], s])
        m, nt = self.tau.size, len(live)
        M = np.ones((rows.size, self._ncoef()))
        # s / (s + 1/tau) written so the elastic row (s = inf) is exactly 1
        M[:, 1:1 + m] = 1.0 / (1.0 + 1.0 / (rows[:, None] * self.tau[None, :]))
        if nt:
            M[:, 1 + m:1 + m + nt] = (1.0 / (1.0 + self._rate[None, :] / rows[:, None])) @ self._w.T
            M[0, 1 + m:1 + m + nt] += self._rest     # past the comb: only the elastic row
        if self.secular:
            M[:, -1] = 1.0 / rows
        self.cond = float(np.linalg.cond(M))
        self._pinv = np.linalg.pinv(M, rcond=defaults.PRONY_RCOND if rcond is None else rcond)

    def _ncoef(self) -> int:
        return 1 + self.tau.size + len(self.tails) + int(self.secular)

    @property
    def window(self) -> tuple:
        """The span of the relaxation times, where the fit is constrained."""
        return float(self.tau.min()), float(self.tau.max())

    def _basis(self, t, kind: str, period=None) -> np.ndarray:
        """Rows of the time functions behind each coefficient."""
        t = np.atleast_1d(np.asarray(t, float))
        m, nt = self.tau.size, len(self.tails)
        e = np.exp(-t[:, None] / self.tau[None, :])
        B = np.zeros((t.size, self._ncoef()))
        cols = slice(1 + m, 1 + m + nt)
        # a tail's comb is exponentials at the rates beyond the grid; its
        # rest has relaxed by any t > 0
        f = np.exp(-t[:, None] * self._rate[None, :]) if nt else None
        if kind == "value":
            B[:, 0] = 1.0
            B[:, 1:1 + m] = e
            if nt:
                B[:, cols] = f @ self._w.T + np.outer(t == 0.0, self._rest)
            if self.secular:
                B[:, -1] = t
            return B
        if kind == "rate":
            B[:, 1:1 + m] = -e / self.tau
            if nt:
                B[:, cols] = -(f * self._rate) @ self._w.T
            if self.secular:
                B[:, -1] = 1.0
            return B
        if self.secular:
            raise ValueError("a creeping response has no periodic steady state")
        # a step repeated every period: sum over n of the relaxation since
        # each, a geometric series in exp(-period / tau)
        geo = 1.0 / (1.0 - np.exp(-float(period) / self.tau))
        far = 1.0 / (1.0 - np.exp(-float(period) * self._rate)) if nt else None
        if kind == "cycle":
            B[:, 1:1 + m] = (e - 1.0) * geo
            if nt:
                B[:, cols] = ((f - 1.0) * far) @ self._w.T - np.outer(t > 0.0, self._rest)
        else:
            B[:, 1:1 + m] = -e / self.tau * geo
            if nt:
                B[:, cols] = -(f * self._rate * far) @ self._w.T
        return B

    def coefficients(self) -> np.ndarray:
        """(n_coef, n_states): the map from a stack to [a0, a_1 .. a_m,
        (one amplitude per tail), (c)]."""
        return self._pinv

    def memory(self, decades: float = None) -> tuple:
        """(tau, C, gone): the fit as exponentials and nothing else
        (laplace.py:232-258).  C maps a stack to [a0, a_j] over tau; a
        tail's comb poles within ``decades`` (``STEPPER_TAIL_DECADES``) of
        the fastest fitted rate are kept as poles, and the rest of the comb
        plus its remainder is ``gone @ stack``, relaxed at any lag > 0.  The
        elastic instant is C summed plus gone."""
        if self.secular:
            raise ValueError("a creeping response is not a sum of exponentials")
        m, nt = self.tau.size, len(self.tails)
        if not nt:
            return self.tau, self._pinv, np.zeros(self._pinv.shape[1])
        decades = defaults.STEPPER_TAIL_DECADES if decades is None else decades
        keep = int(np.searchsorted(self._rate * self.tau.min(), 10.0 ** decades * (1.0 + 1e-9)))
        amp = self._pinv[1 + m:1 + m + nt]
        return (np.concatenate([self.tau, 1.0 / self._rate[:keep]]),
                np.vstack([self._pinv[:1 + m], self._w[:, :keep].T @ amp]),
                (self._w[:, keep:].sum(axis=1) + self._rest) @ amp)

    def weights(self, t) -> np.ndarray:
        """(n_t, n_states): u(t) = weights(t) @ stack for a unit step."""
        return self._basis(t, "value") @ self._pinv

    def rate_weights(self, t) -> np.ndarray:
        """The same for du/dt."""
        return self._basis(t, "rate") @ self._pinv

    def creep_rate_weights(self) -> np.ndarray:
        """(n_states,): the steady creep rate c of a secular fit."""
        if not self.secular:
            raise ValueError("fitted without the secular term: no creep rate")
        return self._pinv[-1]

    def cycle_weights(self, t, period: float) -> np.ndarray:
        """The step repeated every ``period`` for ever: displacement since the
        latest one, at t after it, from the relaxation of all of them."""
        return self._basis(t, "cycle", period) @ self._pinv

    def cycle_rate_weights(self, t, period: float) -> np.ndarray:
        """... and its rate."""
        return self._basis(t, "cycle_rate", period) @ self._pinv

    def relaxed_weights(self) -> np.ndarray:
        """(n_states,): the fit's constant u(inf) (the end of the window)."""
        if self.secular:
            raise ValueError("a creeping response never settles")
        return self._pinv[0]

    def locked_rate_weights(self, t, period: float) -> np.ndarray:
        """(n_t, n_states): velocity per unit slip rate of a fault locked
        between earthquakes every ``period``, measured from steady slip:
        period * sum_n du/dt(t + n period) - u(inf).  Its mean over a cycle
        is minus the elastic response (the back-slip model)."""
        return (float(period) * self.cycle_rate_weights(t, period)
                - self.relaxed_weights()[None, :])


def slug(label) -> str:
    """A state label as it appears in a file name: runs of anything but
    letters, digits, '.', '_' and '-' become one '-' (config.py:47-54)."""
    return re.sub(r"[^A-Za-z0-9._-]+", "-", str(label)).strip("-") or "x"


def _json_value(v):
    """JSON-safe view: non-finite floats as the strings "inf"/"-inf"/"nan"
    (provenance.py:105-128), dataclasses as dicts, tuples as lists."""
    if dataclasses.is_dataclass(v) and not isinstance(v, type):
        return _json_value(dataclasses.asdict(v))
    if isinstance(v, dict):
        return {str(k): _json_value(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_json_value(x) for x in v]
    if isinstance(v, (np.floating, float)):
        v = float(v)
        if math.isnan(v):
            return "nan"
        if math.isinf(v):
            return "inf" if v > 0 else "-inf"
        return v
    if isinstance(v, np.integer):
        return int(v)
    if isinstance(v, np.ndarray):
        return _json_value(v.tolist())
    return v


def _write_json(path, payload) -> None:
    path.write_text(json.dumps(_json_value(payload), indent=1, sort_keys=True,
                               allow_nan=False) + "\n")


def write_run(run_dir, states, solutions, rheology, reports=None, name: str = None):
    """Store solved Laplace states in the run-directory layout of
    runner.py:318-413, the part ``run_stack`` / ``run_fit`` (and the
    reference's own readers) take:

      resolved.json   schema 1; spec.laplace = {rheology, s, faults}; the
                      states as solved (label, materials, s, drag)
      fields_<slug(label)>.npz   one per state: slot name -> (n_dofs, 3)
      report.json     one row per state (label + ``reports[i]``)
      MANIFEST        sha256, size and path of every other file

    ``states``: ``rheology.State`` (from ``laplace_states``), the elastic one
    among them; ``solutions[i]``: {slot: array} of ``states[i]``;
    ``rheology``: region -> rheology (object or dict form).  Refuses labels
    whose file names would collide."""
    d = pathlib.Path(run_dir)
    if len(states) != len(solutions):
        raise ValueError(f"{len(states)} states and {len(solutions)} solutions")
    slugs = [slug(st.label) for st in states]
    if len(set(slugs)) != len(slugs):
        raise ValueError(f"state labels collide as file names: {slugs}")
    if not any(st.s is not None and math.isinf(float(st.s)) for st in states):
        raise ValueError("a Laplace run needs its elastic (s = inf) state")
    d.mkdir(parents=True, exist_ok=True)
    rheo = {k: (dataclasses.asdict(v) if dataclasses.is_dataclass(v) else dict(v))
            for k, v in rheology.items()}
    finite = sorted((float(st.s) for st in states
                     if st.s is not None and math.isfinite(float(st.s))), reverse=True)
    _write_json(d / "resolved.json", {
        "schema": 1,
        "run": {"id": d.name, "name": name or d.name},
        "spec": {"laplace": {"rheology": rheo, "s": finite, "faults": {}}},
        "states": [{"label": st.label, "s": st.s, "drag": dict(st.drag),
                    "materials": {k: dataclasses.asdict(m) for k, m in st.materials.items()}}
                   for st in states],
    })
    rows = []
    for i, (st, sol) in enumerate(zip(states, solutions)):
        np.savez(d / f"fields_{slugs[i]}.npz", **{k: np.asarray(v) for k, v in sol.items()})
        rows.append({"label": st.label, **(dict(reports[i]) if reports else {})})
    _write_json(d / "report.json", {"states": rows})
    lines = []
    for f in sorted(d.rglob("*")):
        if f.is_file() and f.name != "MANIFEST":
            b = f.read_bytes()
            lines.append(f"{hashlib.sha256(b).hexdigest()}  {len(b):>12d}  {f.relative_to(d)}")
    (d / "MANIFEST").write_text("\n".join(lines) + "\n")
    return d


def _run_states(run_dir) -> tuple:
    """(directory, state rows elastic first then by descending s, spec) of a
    stored Laplace run, from its resolved.json (laplace.py:315-330)."""
    d = pathlib.Path(run_dir)
    resolved = json.loads((d / "resolved.json").read_text())
    rows = [r for r in resolved.get("states", []) if r.get("s") is not None]
    if not rows:
        raise ValueError(f"{d.name} is not a Laplace run: none of its states carries a "
                         f"Laplace parameter")
    rows.sort(key=lambda r: -float(r["s"]))
    if not math.isinf(float(rows[0]["s"])):
        raise ValueError(f"{d.name} has no elastic (s = inf) state")
    return d, rows, resolved["spec"]


def run_stack(run_dir) -> tuple:
    """(s, {slot: (1 + len(s), ...) array}) of a stored Laplace run, in the
    order ``Prony`` takes (elastic, then descending s), read from the run's
    own record of its states, never from file names (laplace.py:333-354)."""
    d, rows, _ = _run_states(run_dir)
    stack: dict = {}
    for r in rows:
        with np.load(d / f"fields_{slug(r['label'])}.npz") as z:
            for k in z.files:
                stack.setdefault(k, []).append(z[k])
    return (np.array([float(r["s"]) for r in rows[1:]]),
            {k: np.stack(v) for k, v in stack.items()})


def run_fit(run_dir, **kw) -> Prony:
    """``Prony`` on a stored run's points with the tails of its rheologies
    (laplace.py:357-367); ``kw`` goes to ``Prony`` (``secular``)."""
    from . import rheology

    _, rows, spec = _run_states(run_dir)
    rheo = ((spec.get("laplace") or {}).get("rheology") or {}).values()
    return Prony([float(r["s"]) for r in rows[1:]], tails=rheology.tails(rheo), **kw)


def apply_weights(W, stack):
    """out[i, ...] = sum_k W[i, k] stack[k, ...] on the device: the time
    response of a state stack (first axis = states, in ``Prony`` order; host
    or device), one library DGEMM.  Returns a device tensor."""
    import torch

    from . import _lib
    from ._lib import call, ptr, stream
    _lib.require_gpu()
    dev = _lib.device()
    S = torch.as_tensor(stack, dtype=torch.float64, device=dev)
    W = torch.as_tensor(np.atleast_2d(np.asarray(W, float)), dtype=torch.float64, device=dev)
    ns = S.shape[0]
    if W.shape[1] != ns:
        raise ValueError(f"weights have {W.shape[1]} states, the stack {ns}")
    S2 = S.reshape(ns, -1).contiguous()
    W = W.contiguous()
    out = torch.empty((W.shape[0], S2.shape[1]), dtype=torch.float64, device=dev)
    call("mbem_dgemm_rowmajor", W.shape[0], S2.shape[1], ns, 1.0, ptr(W), ns, ptr(S2),
         S2.shape[1], 0.0, ptr(out), S2.shape[1], stream())
    return out.reshape((W.shape[0],) + tuple(S.shape[1:]))
