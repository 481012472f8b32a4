"""The learned coarse space on the GPU (mbem_amd/learned_coarse.py and the
global columns of mbem_amd/la/coarse.py) against the CPU oracle:

* the ridge subspace loss and its gradient dL/dy of the device (FP32 basis
  and targets, FP64 products) equal oracle.learned.subspace_loss (float64,
  itself pinned by central differences) to 1e-4 relative;
* the two-level preconditioner with m global columns E beside the groups'
  shape functions (basis "rigid") and alone (basis "none") equals the
  oracle's two_level with W = [W_groups | E] to 1e-9 on the 512-triangle
  box, and deflates exactly: r = A w with w in span W gives z = w.
"""
import numpy as np
import pytest
import torch

from oracle import coarse as oc
from oracle import dense as od
from oracle import learned as ol

pytestmark = pytest.mark.gpu


def test_subspace_loss_matches_oracle():
    from mbem_amd import learned as L
    from mbem_amd import learned_coarse as LC
    from test_learned_gpu import _FakeCloud
    cl = _FakeCloud(900, seed=3)
    rng = np.random.default_rng(5)
    y = rng.normal(size=(cl.P, L.N_OUT)).astype(np.float32)
    cl.y.copy_(torch.as_tensor(y))
    V = rng.normal(size=(7, 3 * cl.P)).astype(np.float32)
    lam = 0.05
    dy = torch.empty_like(cl.y)
    loss = LC.subspace_loss_grad(None, cl, torch.as_tensor(V, device=cl.dev), dy, lam=lam)
    want, dwant = ol.subspace_loss(y.astype(np.float64), V.astype(np.float64), lam)
    got = dy.cpu().numpy()
    print(f"loss {loss:.6f} oracle {want:.6f}; |dy - oracle| / |oracle| = "
          f"{np.abs(got - dwant).max() / np.abs(dwant).max():.1e}", flush=True)
    assert abs(loss - want) <= 1e-4 * want
    assert np.abs(got - dwant).max() <= 1e-4 * np.abs(dwant).max()
    assert not np.any(got[:, 3 * L.RANK:])


@pytest.fixture(scope="module")
def box():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.la.preconditioner import BlockGaussSeidel
    from mbem_amd.model import generate_system
    model, eps, _ = box_fault_model(8, fault_cells=(6, 10))
    system = generate_system(model)
    asm = DenseBackend().assemble(system, eps)
    M = BlockGaussSeidel(asm, max_dense=0, bj_chunk=150, grouping="region")
    A, b, _ = od.assemble(system, eps)
    return dict(asm=asm, M=M, A=A, b=b)


@pytest.mark.parametrize("basis", ["none", "rigid"])
def test_global_columns_match_oracle(box, basis):
    from mbem_amd.la.coarse import CoarseSpace, TwoLevel, dof_geometry
    asm, M, A = box["asm"], box["M"], box["A"]
    n = A.shape[0]
    rng = np.random.default_rng(11)
    # smooth global columns: low-order polynomials of the unknowns' points
    pts, kind, comp = dof_geometry(asm)
    s = pts / np.abs(pts).max()
    E = np.stack([np.ones(n), s[:, 0], s[:, 1] * s[:, 2], np.sin(3 * s[:, 0]) + comp,
                  rng.standard_normal(n)], axis=1)
    cs = CoarseSpace(asm, M.chunk_dofs, basis=basis,
                     extra=torch.as_tensor(E, device=asm.A_dev.device)).finish()
    assert cs.m == E.shape[1]
    W = E if basis == "none" else np.hstack([oc.coarse_basis(M.chunk_dofs, pts, kind, comp,
                                                             basis), E])
    assert cs.K == W.shape[1]
    P = TwoLevel(M, cs)
    Po = oc.two_level(A, W, oc.block_jacobi(A, M.chunk_dofs))
    for r in (box["b"], rng.standard_normal(n)):
        z = P(torch.as_tensor(r, device=asm.A_dev.device)).cpu().numpy()
        zo = Po(r)
        err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
        print(f"{basis} + {cs.m} global: rank {cs.Kp} of {cs.K}, |z - z_oracle| / |z| = {err:.1e}",
              flush=True)
        assert err < 1e-9
    w = W @ rng.standard_normal(W.shape[1])
    z = P(torch.as_tensor(A @ w, device=asm.A_dev.device)).cpu().numpy()
    assert np.linalg.norm(z - w) < 1e-9 * np.linalg.norm(w)
    with pytest.raises(ValueError):
        CoarseSpace(asm, M.chunk_dofs, basis=basis,
                    extra=torch.zeros((n - 1, 2), dtype=torch.float64, device=asm.A_dev.device))
