"""The near-face readout on the GPU (mbem_amd.evaluate with ``near=``, the
recovered density's near pairs on the nodal kernels of csrc/nodal.hip)
against the oracle's (oracle/evaluate.py ``near=``, pinned by Somigliana's
identity in tests/test_near_readout.py), on the same densities:

* the exact densities of u = c + E x on the 8-cell box, at 0.1 and 0.3 h
  below the top face and in the interior: equal to the oracle to 1e-10 of
  the field's size, and within the oracle test's bars of the exact field;
* the solved box with its dipping fault (a solution with every density
  live, the fault untouched by the recovery), at points 0.1-0.5 h under the
  top face and beside a wall: equal to the oracle to 1e-10, stress (elastic
  and with ``parts``) and displacement.
"""
import numpy as np
import pytest

import _bvp as B
from oracle import evaluate as oe

pytestmark = pytest.mark.gpu


def test_exact_density_near_readout_gpu():
    from mbem_amd import evaluate as ev
    from mbem_amd.recover import NearField
    model, region, sol, mat = B.exact_density_model(0.25, n_cells=8)
    h = 5.0
    eps = 0.125 * h
    near = NearField(model)
    rng = np.random.default_rng(1)
    xy = rng.uniform(-12.0, 12.0, (6, 2))
    se = B.exact_sigma(mat)
    for depth in (0.1, 0.3, 2.0):
        q = np.column_stack([xy, np.full(6, -depth * h)])
        print(f"[near readout] exact densities, depth {depth} h", flush=True)
        s = ev.evaluate_stress(model, region, sol, q, eps, near=near)
        u = ev.evaluate_displacement(model, region, sol, q, eps, near=near)
        so = oe.evaluate_stress(model, region, sol, q, eps, near=near)
        uo = oe.evaluate_displacement(model, region, sol, q, eps, near=near)
        ue = B.RIGID_C + q @ B.STRAIN.T
        ds = np.abs(s - so).max() / np.linalg.norm(se)
        du = np.abs(u - uo).max() / np.abs(ue).max()
        es = np.abs(s - se).max() / np.linalg.norm(se)
        eu = np.abs(u - ue).max() / np.abs(ue).max()
        print(f"  vs oracle {ds:.1e} / {du:.1e}; vs exact {es:.1e} / {eu:.1e}", flush=True)
        assert ds < 1e-10 and du < 1e-10
        assert es < 5e-3 and eu < 5e-3


def test_solved_box_near_readout_gpu():
    from mbem_amd import evaluate as ev
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    from mbem_amd.recover import NearField
    model, eps, _ = box_fault_model(8, fault_cells=(6, 10))
    system = generate_system(model)
    asm = DenseBackend().assemble(system, eps)
    sol = asm.solve(rtol=1e-12, precond_max_dense=0)
    region = model.regions[0]
    near = NearField(model)
    rng = np.random.default_rng(3)
    h = 5.0
    q = np.concatenate([
        np.column_stack([rng.uniform(-15, 15, (8, 2)), -rng.uniform(0.1, 0.5, 8) * h]),
        np.column_stack([np.full(4, 20.0 - 0.2 * h), rng.uniform(-10, 10, 4),
                         rng.uniform(-15, -5, 4)])])
    print("[near readout] solved box", flush=True)
    s = ev.evaluate_stress(model, region, sol, q, eps, near=near)
    so = oe.evaluate_stress(model, region, sol, q, eps, near=near)
    sp, ep = ev.evaluate_stress(model, region, sol, q, eps, near=near, parts=True)
    u = ev.evaluate_displacement(model, region, sol, q, eps, near=near)
    uo = oe.evaluate_displacement(model, region, sol, q, eps, near=near)
    ds = np.abs(s - so).max() / np.abs(so).max()
    du = np.abs(u - uo).max() / np.abs(uo).max()
    print(f"  stress vs oracle {ds:.1e}, displacement {du:.1e}", flush=True)
    assert ds < 1e-10 and du < 1e-10
    assert np.abs(sp - s).max() < 1e-12 * np.abs(s).max()
    # elastic = total - sum sigma C:eps*: parts gives the two pieces
    s_tot = ev.evaluate_stress(model, region, sol, q, eps, near=near, subtract_anelastic=False)
    assert np.abs((s_tot + ep) - s).max() < 1e-10 * np.abs(s).max()
