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
"""
from __future__ import annotations

import math

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


class Prony:
    """The fit for one set of finite collocation points ``s``; a stack is
    ordered [elastic, s[0], s[1], ...] (laplace.py:119-312).

    ``secular`` adds the creep term c t and drops the slowest pole (the two
    are all but collinear over the window).  Power-law tails (bodies with no
    fastest relaxation) need rheologies this build does not have, and are
    refused."""

    def __init__(self, s, secular: bool = False, rcond: float = None, tails=()):
        s = np.asarray(s, float)
        if s.ndim != 1 or s.size < 2 or np.any(s <= 0) or np.any(~np.isfinite(s)):
            raise ValueError("s must be at least two finite positive values")
        if np.unique(s).size != s.size:
            raise ValueError("collocation points must be distinct")
        if tails:
            raise NotImplementedError("power-law tails (Andrade, extended Burgers) need "
                                      "rheologies this build does not have")
        self.s = s
        self.secular = bool(secular)
        self.tails = ()
        tau = np.sort(1.0 / s)
        self.tau = tau[:-1] if self.secular else tau
        rows = np.concatenate([[np.inf], s])
        M = np.ones((rows.size, self._ncoef()))
        # s / (s + 1/tau) written so the elastic row (s = inf) is exactly 1
        M[:, 1:1 + self.tau.size] = 1.0 / (1.0 + 1.0 / (rows[:, None] * self.tau[None, :]))
        if self.secular:
            M[:, -1] = 1.0 / rows
        self.cond = float(np.linalg.cond(M))
        self._pinv = np.linalg.pinv(M, rcond=defaults.PRONY_RCOND if rcond is None else rcond)

    def _ncoef(self) -> int:
        return 1 + self.tau.size + int(self.secular)

    @property
    def window(self) -> tuple:
        """The span of the relaxation times, where the fit is constrained."""
        return float(self.tau.min()), float(self.tau.max())

    def _basis(self, t, kind: str, period=None) -> np.ndarray:
        """Rows of the time functions behind each coefficient."""
        t = np.atleast_1d(np.asarray(t, float))
        m = self.tau.size
        e = np.exp(-t[:, None] / self.tau[None, :])
        B = np.zeros((t.size, self._ncoef()))
        if kind == "value":
            B[:, 0] = 1.0
            B[:, 1:1 + m] = e
            if self.secular:
                B[:, -1] = t
            return B
        if kind == "rate":
            B[:, 1:1 + m] = -e / self.tau
            if self.secular:
                B[:, -1] = 1.0
            return B
        if self.secular:
            raise ValueError("a creeping response has no periodic steady state")
        # a step repeated every period: sum over n of the relaxation since
        # each, a geometric series in exp(-period / tau)
        geo = 1.0 / (1.0 - np.exp(-float(period) / self.tau))
        if kind == "cycle":
            B[:, 1:1 + m] = (e - 1.0) * geo
        else:
            B[:, 1:1 + m] = -e / self.tau * geo
        return B

    def coefficients(self) -> np.ndarray:
        """(n_coef, n_states): the map from a stack to [a0, a_1 .. a_m, (c)]."""
        return self._pinv

    def memory(self, decades: float = None) -> tuple:
        """(tau, C, gone): the fit as exponentials -- C maps a stack to
        [a0, a_j] over tau, and nothing has relaxed at once (no tails)."""
        if self.secular:
            raise ValueError("a creeping response is not a sum of exponentials")
        return self.tau, self._pinv, np.zeros(self._pinv.shape[1])

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
