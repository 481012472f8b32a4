"""GPU parity of the dense LU solves (the reference's LU path,
backends/dense.py:532-571 and la/lu.py): rocSOLVER's factors, then the
blocked triangular solves (one launch per 64-row diagonal block, inverses of
the diagonal blocks formed once) for A x = b and A^T x = b.

Bars: against numpy's float64 solve of the same matrix, relative 2-norm
within 1e-12 times the 1-norm condition number; against rocSOLVER's own
getrs on the same factors within 1e-13 times it.  Sizes cover one partial
block, exactly one block, one block plus one row, a few blocks with a
partial last one, and a matrix whose pivoting permutes most rows; the
pivot order and the diagonal blocks' inverses are checked on their own.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _factor(A):
    from mbem_amd._lib import call, ptr, stream
    from mbem_amd.backends.dense import _LUFactors
    n = A.shape[0]
    dev = torch.device("cuda")
    LU = torch.as_tensor(A, device=dev).clone()
    ipiv = torch.empty(n, dtype=torch.int64, device=dev)
    info = torch.zeros(1, dtype=torch.int64, device=dev)
    call("mbem_lu_factor", ptr(LU), n, ptr(ipiv), ptr(info), stream())
    assert int(info.item()) == 0
    return _LUFactors(LU, ipiv, n)


@pytest.mark.parametrize("n", [1, 37, 64, 65, 300, 1000])
@pytest.mark.parametrize("kind", ["dominant", "permuting"])
def test_blocked_solves_match_numpy_and_getrs(n, kind):
    from mbem_amd._lib import call, ptr, stream
    rng = np.random.default_rng(n + (7 if kind == "permuting" else 0))
    A = rng.standard_normal((n, n))
    if kind == "dominant":
        A += n * np.eye(n)
    else:
        # large entries below the diagonal: partial pivoting swaps most rows
        A = np.tril(A, -1) * 50.0 + np.triu(A)
    b = rng.standard_normal(n)
    cond = np.linalg.cond(A, 1)
    fac = _factor(A)
    dev = fac.LU.device
    for adjoint in (False, True):
        x = fac.solve_into(torch.as_tensor(b, device=dev), torch.empty(n, dtype=torch.float64,
                                                                       device=dev),
                           adjoint=adjoint).cpu().numpy()
        ref = np.linalg.solve(A.T if adjoint else A, b)
        e_np = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        g = torch.as_tensor(b, device=dev).clone()
        call("mbem_lu_solve_adjoint" if adjoint else "mbem_lu_solve", ptr(fac.LU), n,
             ptr(fac.ipiv), ptr(g), 1, stream())
        e_rs = np.linalg.norm(x - g.cpu().numpy()) / np.linalg.norm(ref)
        print(f"n={n} {kind} adjoint={adjoint}: vs numpy {e_np:.1e}, vs getrs {e_rs:.1e}, "
              f"cond {cond:.1e}", flush=True)
        assert e_np < 1e-12 * cond and e_rs < 1e-13 * cond
    # in place (b and x the same vector)
    v = torch.as_tensor(b, device=dev).clone()
    fac.solve_into(v, v)
    assert np.linalg.norm(v.cpu().numpy() - np.linalg.solve(A, b)) <= \
        1e-12 * cond * np.linalg.norm(np.linalg.solve(A, b))


def test_pivot_order_and_block_inverses():
    """perm reproduces rocSOLVER's interchanges (P^T A^T = L U rows), and
    each stored block inverse times its block is the identity."""
    n = 150
    rng = np.random.default_rng(3)
    A = np.tril(rng.standard_normal((n, n)), -1) * 50.0 + np.triu(rng.standard_normal((n, n)))
    fac = _factor(A)
    # column-major view of the row-major buffer: M = A^T = P L U
    M = fac.LU.cpu().numpy().T
    L = np.tril(M, -1) + np.eye(n)
    U = np.triu(M)
    perm = fac.perm.cpu().numpy()
    assert sorted(perm.tolist()) == list(range(n))
    assert np.allclose(A.T[perm], L @ U, rtol=0, atol=1e-11 * np.abs(A).max())
    inv_l = fac.inv_l.cpu().numpy().reshape(-1, 64, 64)
    inv_u = fac.inv_u.cpu().numpy().reshape(-1, 64, 64)
    for k in range(inv_l.shape[0]):
        k0, k1 = 64 * k, min(64 * k + 64, n)
        nb = k1 - k0
        assert np.allclose(inv_l[k][:nb, :nb] @ L[k0:k1, k0:k1], np.eye(nb), atol=1e-12)
        assert np.allclose(inv_u[k][:nb, :nb] @ U[k0:k1, k0:k1], np.eye(nb), atol=1e-10)
        assert np.array_equal(inv_l[k][nb:, nb:], np.eye(64 - nb))


def _factor32(A):
    from mbem_amd._lib import call, ptr, stream
    from mbem_amd.backends.dense import _LUFactors
    n = A.shape[0]
    dev = torch.device("cuda")
    LU = torch.as_tensor(A, device=dev).to(torch.float32).contiguous()
    ipiv = torch.empty(n, dtype=torch.int64, device=dev)
    info = torch.zeros(1, dtype=torch.int64, device=dev)
    call("mbem_lu_factor_f32", ptr(LU), n, ptr(ipiv), ptr(info), stream())
    assert int(info.item()) == 0
    return _LUFactors(LU, ipiv, n)


@pytest.mark.parametrize("n", [65, 300, 1000])
def test_single_precision_factors_and_refinement(n):
    """FP32 factors (sgetrf + the blocked solves reading FP32 factors with
    FP64 vectors): one solve is within single precision times the condition
    number of numpy's FP64 solve, both directions; refined against the FP64
    matrix (la/lu.py:198-239) it stops within 3 corrections with the
    predicted remainder below 1e-10, and the actual error is below that."""
    from mbem_amd.backends.dense import refine
    from mbem_amd.la.solver import gemv
    rng = np.random.default_rng(n + 40)
    A = rng.standard_normal((n, n)) + 0.5 * np.sqrt(n) * np.eye(n)
    b = rng.standard_normal(n)
    fac = _factor32(A)
    dev = torch.device("cuda")
    bd = torch.as_tensor(b, device=dev)
    ref = np.linalg.solve(A, b)
    kappa = np.linalg.cond(A, 1)
    x = torch.empty_like(bd)
    fac.solve_into(bd, x)
    e1 = np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref)
    fac.solve_into(bd, x, adjoint=True)
    et = np.linalg.norm(x.cpu().numpy() - np.linalg.solve(A.T, b)) / np.linalg.norm(ref)
    Ad = torch.as_tensor(A, device=dev)
    xr, rep = refine(fac.solve_into, lambda v, out: gemv(Ad, v, out), bd)
    er = np.linalg.norm(xr.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"n {n}: cond {kappa:.1e}; FP32 factors {e1:.1e} (A^T {et:.1e}); refined {er:.1e} "
          f"after {rep['passes']} corrections (left {rep['left']:.1e})", flush=True)
    assert e1 < 1e-6 * kappa and et < 1e-5 * kappa
    assert rep["done"] and rep["passes"] <= 3 and rep["left"] <= 1e-10 and er <= 1e-10


@pytest.mark.parametrize("how", ["single", "gpu"])
def test_dense_lu_single_precision_matches_double(how):
    """lu="single" / "gpu" on the assembled C1-size box: the refined solution
    within 1e-10 of the FP64 LU's, the report carrying the refinement, the
    condition estimate (the same estimator on the FP32 factors) within a
    factor 2 of gecon's on the FP64 ones.  And the fallback: a refinement that
    does not contract (forced) warns and re-factors in double."""
    import warnings

    import mbem_amd.backends.dense as dense
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    model, eps, _ = box_fault_model(16)
    system = generate_system(model)
    ref = AssembledDense(system, eps)
    x64 = ref.solve(method="lu", return_device=True)
    c64 = ref.report.cond_estimate
    asm = AssembledDense(system, eps, lu=how)
    x = asm.solve(method="lu", return_device=True)
    rep = asm.report
    e = float(torch.linalg.vector_norm(x - x64) / torch.linalg.vector_norm(x64))
    print(f"{how}: {rep.lu}; vs double LU {e:.1e}; cond {rep.cond_estimate:.2e} vs {c64:.2e}",
          flush=True)
    assert rep.lu["how"] == how and 1 <= rep.lu["passes"] <= 4 and rep.lu["left"] <= 1e-10
    assert e < 1e-10 and rep.residual < 1e-12
    assert 0.5 <= rep.cond_estimate / c64 <= 2.0
    real = dense.refine
    try:
        dense.refine = lambda *a, **k: (real(*a, **k)[0], {"done": False, "passes": 4,
                                                            "correction": 1.0, "left": 1.0,
                                                            "rate": 0.5})
        asm.release()
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            x2 = asm.solve(method="lu", return_device=True)
        assert any("did not contract" in str(m.message) for m in w)
    finally:
        dense.refine = real
    e2 = float(torch.linalg.vector_norm(x2 - x64) / torch.linalg.vector_norm(x64))
    assert asm.report.lu.get("fallback") is True and e2 < 1e-12
