"""Real Laplace collocation points (laplace.py:61-85).

For step data the transformed solution s u^(s) is the elastic solution at the
moduli of s; a run solves one elastic state per point of ``s_grid``.  The fit
from the state stack back to time (``Prony``) is outside the dense hot path.
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
