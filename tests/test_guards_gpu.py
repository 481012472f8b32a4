"""Failure guards and lifetime on the GPU.

  * FGMRES stops on stagnation (solver.py:214-245): on the cyclic shift,
    whose Krylov residual does not move for n - 1 steps, it stops after
    ``stagnation_window`` iterations without a factor-``stagnation_factor``
    drop, reports ``stagnated`` and not ``converged``, with the iteration
    count of the oracle's restatement of the reference's loop;
  * a numpy-world caller (numpy b, numpy matrix, callables on numpy vectors,
    as the reference's own callers pass) gets a numpy solution equal to the
    device-tensor call's, and a complex right-hand side raises TypeError;
  * the dense LU warns when its condition estimate passes the threshold;
  * the un-anchored refusal holds through ``DenseBackend.assemble``;
  * a dropped assembly frees its device memory at once: after ``del`` (no
    ``gc.collect``) a weak reference to it is dead and the allocated bytes
    are back to where they were, after an FGMRES solve, an LU solve and a
    mixed-precision solve of it.
"""
import gc
import warnings
import weakref

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shift(n):
    P = np.zeros((n, n))
    P[np.arange(1, n), np.arange(n - 1)] = 1.0
    P[0, n - 1] = 1.0
    return P


def test_stagnation_stop_matches_oracle():
    import torch
    from mbem_amd.la.solver import fgmres
    from oracle import dense as od
    n, window = 64, 12
    P = _shift(n)
    b = np.zeros(n)
    b[0] = 1.0
    x, rep = fgmres(torch.as_tensor(P, device="cuda"), torch.as_tensor(b, device="cuda"),
                    rtol=1e-10, restart=n, maxiter=n, stagnation_window=window)
    _, ref = od.fgmres(lambda v: P @ v, b, rtol=1e-10, restart=n, maxiter=n,
                       stagnation_window=window)
    print(rep, ref["iterations"])
    assert rep.stagnated and not rep.converged
    assert rep.iterations == ref["iterations"] == window + 1
    # without the stop the same system converges in n steps
    x, rep = fgmres(torch.as_tensor(P, device="cuda"), torch.as_tensor(b, device="cuda"),
                    rtol=1e-10, restart=n, maxiter=n, stagnation_window=10 * n)
    assert rep.converged and rep.iterations == n and not rep.stagnated


def test_numpy_callers_get_numpy():
    import torch
    from mbem_amd.la.solver import fgmres
    rng = np.random.default_rng(1)
    n = 200
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    d = 1.0 / np.diag(A)
    ref = np.linalg.solve(A, b)
    for Aop, M in ((A, None), (lambda v: A @ v, lambda v: d * v),
                   (type("Op", (), {"matvec": lambda s, v: A @ v})(), None)):
        x, rep = fgmres(Aop, b, M=M, rtol=1e-12)
        assert isinstance(x, np.ndarray) and rep.converged
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-10
    xd, rep = fgmres(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"),
                     rtol=1e-12)
    assert isinstance(xd, torch.Tensor) and xd.is_cuda
    assert np.allclose(xd.cpu().numpy(), ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    with pytest.raises(TypeError, match="real-only"):
        fgmres(A, b + 1j)


@pytest.fixture(scope="module")
def small():
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    model, eps, _ = box_fault_model(8, fault_cells=(6, 10))
    return generate_system(model), eps


def test_lu_condition_warning(small, monkeypatch):
    from mbem_amd import defaults
    from mbem_amd.backends.dense import DenseBackend
    system, eps = small
    asm = DenseBackend().assemble(system, eps)
    with warnings.catch_warnings(record=True) as seen:
        warnings.simplefilter("always")
        asm.solve(method="lu")                      # well conditioned: silent
    assert not [w for w in seen if "condition estimate" in str(w.message)]
    cond = asm.report.cond_estimate
    assert 1.0 < cond < defaults.COND_WARN_THRESHOLD
    asm.release()
    monkeypatch.setattr(defaults, "COND_WARN_THRESHOLD", cond / 2)
    with pytest.warns(UserWarning, match="condition estimate"):
        asm.solve(method="lu")


def test_unanchored_refused_on_device():
    from mbem_amd.backends.dense import DenseBackend
    from test_guards import _free_box
    with pytest.raises(ValueError, match="rigid translations"):
        DenseBackend().assemble(_free_box(), 1.0)


def test_dropped_assembly_frees_its_memory(small):
    import torch
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.la.mixed import solve_mixed
    system, eps = small

    def use():
        asm = DenseBackend().assemble(system, eps)
        asm.solve(rtol=1e-12, precond_max_dense=0)
        asm.solve(method="lu")
        x, _ = solve_mixed(asm, rtol=1e-12)
        del x
        return asm

    # a first round lets the libraries set up whatever they keep for the
    # process (workspaces), so the second measures the assembly alone
    del_me = use()
    del del_me
    gc.collect()
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    gc.disable()                  # only reference counting may free it
    try:
        asm = use()
        assert torch.cuda.memory_allocated() > base
        ref = weakref.ref(asm)
        del asm
        torch.cuda.synchronize()
        assert ref() is None, "the assembly is kept alive by a reference cycle"
        assert torch.cuda.memory_allocated() <= base, (torch.cuda.memory_allocated(), base)
    finally:
        gc.enable()
