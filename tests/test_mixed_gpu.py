"""Mixed-precision operator and solve (north-star pattern 1) on the GPU, on
the 2,048-triangle box with its dipping fault.

  * the near-field blocks are copied exactly (FP64) and the low copy is zero
    there, so (F + N) equals A exactly on the near blocks;
  * the mixed matvec's error against the FP64 matvec is of the storage's
    rounding: below 2e-3 relative (BF16 far field) and 2e-7 (FP32);
  * the nested FGMRES (FP64 outer on the exact operator) reaches the FP64
    true residual 1e-12 and the oracle's LU solution to 1e-9, with and
    without the recycled space across inner solves; the recycled pairs keep
    C^T C = I to 1e-10 and op U = C to 1e-5 (the mixed operator accumulates in
    FP32, so re-applying it to combinations of earlier inputs differs at FP32
    rounding).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    from oracle import dense as od
    model, eps, _ = box_fault_model(16)
    system = generate_system(model)
    asm = DenseBackend().assemble(system, eps)
    A, b, _ = od.assemble(system, eps)
    return dict(asm=asm, A=A, b=b, x_lu=od.lu_solve(A, b))


@pytest.mark.parametrize("storage,tol", [("bf16", 2e-3), ("fp32", 2e-7)])
def test_mixed_matvec(c1, storage, tol):
    import torch
    from mbem_amd.la.mixed import MixedOperator
    asm = c1["asm"]
    op = MixedOperator(asm, storage=storage, k_near=2.0)
    vals = op.vals.view(-1, 3, 3).cpu().numpy()
    A = asm.A          # the device's own FP64 matrix: the near copy is exact
    for i in range(0, op.nbr, 97):
        for b in range(op.rowptr_h[i], op.rowptr_h[i + 1]):
            j = op.col_h[b]
            assert np.array_equal(vals[b], A[3 * i:3 * i + 3, 3 * j:3 * j + 3])
    rng = np.random.default_rng(0)
    v = rng.standard_normal(A.shape[0])
    y = op.matvec(torch.as_tensor(v, device=asm.A_dev.device)).cpu().numpy()
    ref = A @ v
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"{storage}: near blocks {op.nnzb}, matvec rel err {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("recycle", [0, 60])
def test_mixed_solve(c1, recycle):
    from mbem_amd.la.mixed import solve_mixed
    asm = c1["asm"]
    x, rep = solve_mixed(asm, rtol=1e-12, storage="bf16", recycle=recycle)
    xh = x.cpu().numpy()
    err = np.linalg.norm(xh - c1["x_lu"]) / np.linalg.norm(c1["x_lu"])
    print(f"{rep}; rel err vs oracle LU {err:.2e}")
    assert rep.converged and rep.outer.true_relres < 1e-12
    assert err < 1e-9


def test_recycled_pairs_stay_consistent(c1):
    """After a few inner solves the recycled (U, C) satisfy op U = C, C^T C = I."""
    import torch
    from mbem_amd.la.mixed import MixedOperator, _RecycledInner
    from mbem_amd.la.preconditioner import BlockGaussSeidel
    asm = c1["asm"]
    op = MixedOperator(asm, storage="bf16")
    M = BlockGaussSeidel(asm, max_dense=0, bj_chunk=1500, grouping="region", inverse_dtype="f32",
                         coupling=op)
    inner = _RecycledInner(op, M, 5e-2, 40, op.n, asm.A_dev.device, kmax=30)
    rng = np.random.default_rng(1)
    for _ in range(4):
        r = torch.as_tensor(rng.standard_normal(op.n), device=asm.A_dev.device)
        z = torch.empty_like(r)
        inner.apply_into(r, z)
    k = inner.k
    assert k > 0
    U, C = inner.U[:k], inner.C[:k]
    OU = torch.stack([op.matvec(U[i].contiguous()) for i in range(k)])
    e1 = float((OU - C).norm() / C.norm())
    e2 = float((C @ C.T - torch.eye(k, dtype=C.dtype, device=C.device)).abs().max())
    print(f"recycled k={k}: |op U - C|/|C| {e1:.2e}, |C^T C - I| {e2:.2e}, "
          f"inner iterations {inner.iterations}")
    assert e1 < 1e-5 and e2 < 1e-10
