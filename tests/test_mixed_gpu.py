"""Mixed-precision operator and solve (north-star pattern 1) on the GPU, on
the 2,048-triangle box with its dipping fault.

  * the near-field blocks are copied exactly (FP64) and the low copy is zero
    there, so (F + N) equals A exactly on the near blocks;
  * the mixed matvec's error against the FP64 matvec is of the storage's
    rounding: below 2e-3 relative (BF16 far field) and 2e-7 (FP32);
  * the nested FGMRES (FP64 outer on the exact operator) reaches the FP64
    true residual 1e-12 and the oracle's LU solution to 1e-9.
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


def test_mixed_solve(c1):
    from mbem_amd.la.mixed import solve_mixed
    asm = c1["asm"]
    x, rep = solve_mixed(asm, rtol=1e-12, storage="bf16")
    xh = x.cpu().numpy()
    err = np.linalg.norm(xh - c1["x_lu"]) / np.linalg.norm(c1["x_lu"])
    print(f"{rep}; rel err vs oracle LU {err:.2e}")
    assert rep.converged and rep.outer.true_relres < 1e-12
    assert err < 1e-9
