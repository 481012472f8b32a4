"""The near-face readout (evaluate.py:410-472 with recover.NearField) pinned
by Somigliana's identity on the CPU oracle (oracle/evaluate.py ``near=``).

With the EXACT densities of u = c + E x on a closed box (tests/_bvp.py), the
representation formula must return the field inside.  A point within a
fraction of an element of a face resolves the piecewise-constant density's
edge steps: at 0.1 h below the top face the stress errs by ~15 % of |C:E|
without the near-face readout.  Read from the recovered continuous density
at the narrowed mollification the near pairs must bring it under 5e-3
(measured 1.1e-3 at 0.1 h and 1.5e-3 at 0.3 h, displacement 1.2e-3), and
two plausible mistakes must not: the element values copied to every node
instead of recovered, and the narrowed width left at the solve's.
"""
import numpy as np
import pytest

import _bvp as B
from oracle import evaluate as oe

NC = 8


@pytest.fixture(scope="module")
def case():
    from mbem_amd.recover import NearField
    model, region, sol, mat = B.exact_density_model(0.25, n_cells=NC)
    h = 40.0 / NC
    rng = np.random.default_rng(1)
    xy = rng.uniform(-12.0, 12.0, (6, 2))
    return dict(model=model, region=region, sol=sol, mat=mat, h=h, eps=0.125 * h, xy=xy,
                near=NearField(model))


def _errors(c, depth, near):
    q = np.column_stack([c["xy"], np.full(len(c["xy"]), -depth * c["h"])])
    se = B.exact_sigma(c["mat"])
    ue = B.RIGID_C + q @ B.STRAIN.T
    s = oe.evaluate_stress(c["model"], c["region"], c["sol"], q, c["eps"], near=near)
    u = oe.evaluate_displacement(c["model"], c["region"], c["sol"], q, c["eps"], near=near)
    return (np.abs(s - se).max() / np.linalg.norm(se), np.abs(u - ue).max() / np.abs(ue).max())


@pytest.mark.parametrize("depth", [0.1, 0.3])
def test_near_readout_recovers_the_field_below_a_face(case, depth):
    e_plain, _ = _errors(case, depth, None)
    e_near, d_near = _errors(case, depth, case["near"])
    print(f"depth {depth} h: stress {e_plain:.2e} -> {e_near:.2e}, displacement {d_near:.2e}")
    assert e_plain > 5e-2
    assert e_near < 5e-3 and d_near < 5e-3


class _Copied:
    """A 'recovery' that copies each element's value to its six nodes."""

    def __init__(self, near):
        self.eps_factor, self.radius_h = near.eps_factor, near.radius_h
        self._near = near

    def recovers(self, p):
        return self._near.recovers(p)

    def recover(self, p, values):
        return np.repeat(np.asarray(values, float).reshape(-1, 3), 6, axis=0)


class _Wide(_Copied):
    """The recovered density, but read at the solve's width."""

    def __init__(self, near):
        super().__init__(near)
        self.eps_factor = 1.0

    def recover(self, p, values):
        return self._near.recover(p, values)


@pytest.mark.parametrize("bad", [_Copied, _Wide])
def test_tripwires(case, bad):
    e, _ = _errors(case, 0.1, bad(case["near"]))
    print(f"{bad.__name__}: stress error {e:.2e}")
    assert e > 5e-3
