"""Two-level block-Jacobi on the GPU (mbem_amd/la/coarse.py) against the
oracle's restatement (oracle/coarse.py), on the 512-triangle box:

* the coarse image Ct = (A W)^T of mbem_coarse_image equals numpy's A W
  with the oracle's W to 1e-13 of its largest entry;
* one application z = P r equals the oracle's (exact least squares instead
  of the device's orthogonalised basis, whose scaled Gram matrix is
  regularised by 1e-13 I) to 1e-9, for each basis;
* the deflation identity holds on the device: r = A W y gives z = W y;
* FGMRES with the coarse correction converges to the oracle's LU solution
  in fewer iterations than block-Jacobi alone, and the mixed-precision
  solve with it still reaches the LU solution to 1e-9.
"""
import numpy as np
import pytest

from oracle import coarse as oc
from oracle import dense as od

pytestmark = pytest.mark.gpu


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
    return dict(asm=asm, M=M, A=A, b=b, x_lu=od.lu_solve(A, b))


def _geometry(asm):
    from mbem_amd.la.coarse import dof_geometry
    return dof_geometry(asm)


@pytest.mark.parametrize("basis", ["const", "rigid", "linear"])
def test_apply_matches_oracle(box, basis):
    import torch
    from mbem_amd import _lib
    from mbem_amd.la.coarse import CoarseSpace, TwoLevel
    asm, M, A = box["asm"], box["M"], box["A"]
    pts, kind, comp = _geometry(asm)
    W = oc.coarse_basis(M.chunk_dofs, pts, kind, comp, basis)
    cs = CoarseSpace(asm, M.chunk_dofs, basis=basis).finish()
    # the coarse image itself, recomputed into a fresh buffer
    Ct = torch.empty((cs.K, cs.n), dtype=torch.float64, device=asm.A_dev.device)
    gptr = np.concatenate([[0], np.cumsum(np.bincount(cs.grp.cpu().numpy(), minlength=cs.G))])
    _lib.call("mbem_coarse_image", _lib.ptr(asm.A_dev), asm.A_dev.stride(0), cs.n,
              _lib.ptr(cs.perm), _lib.ptr(cs.Wp),
              _lib.ptr(torch.as_tensor(gptr, device=Ct.device)), cs.G, cs.q, _lib.ptr(Ct), cs.n,
              _lib.stream())
    C = A @ W
    assert np.abs(Ct.cpu().numpy().T - C).max() < 1e-13 * np.abs(C).max()
    P = TwoLevel(M, cs)
    Po = oc.two_level(A, W, oc.block_jacobi(A, M.chunk_dofs))
    rng = np.random.default_rng(2)
    for r in (box["b"], rng.standard_normal(A.shape[0])):
        z = P(torch.as_tensor(r, device=asm.A_dev.device)).cpu().numpy()
        zo = Po(r)
        err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
        print(f"{basis}: rank {cs.Kp} of {cs.K}, |z - z_oracle| / |z| = {err:.1e}")
        assert err < 1e-9
    # deflation identity: r = A w, w in span W  ->  z = w
    w = W @ rng.standard_normal(W.shape[1])
    z = P(torch.as_tensor(A @ w, device=asm.A_dev.device)).cpu().numpy()
    assert np.linalg.norm(z - w) < 1e-9 * np.linalg.norm(w)


def test_fgmres_with_coarse_correction(box):
    asm = box["asm"]
    asm.release()
    asm.solve(rtol=1e-12, precond_max_dense=0, precond_bj_chunk=150, precond_grouping="region")
    it_bj = asm.report.iterations
    asm.release()
    x = asm.solve(rtol=1e-12, precond_max_dense=0, precond_bj_chunk=150, precond_grouping="region",
                  precond_coarse="rigid", return_device=True).cpu().numpy()
    rep = asm.report
    err = np.linalg.norm(x - box["x_lu"]) / np.linalg.norm(box["x_lu"])
    print(f"block-Jacobi {it_bj} iterations, two-level {rep.iterations}; err {err:.1e}; "
          f"{rep.precond_summary['coarse']}")
    assert rep.converged and err < 1e-9
    assert rep.iterations < it_bj
    asm.release()


@pytest.mark.parametrize("storage", ["int8", "bf16"])
def test_mixed_solve_with_coarse_correction(box, storage):
    from mbem_amd.la.mixed import solve_mixed
    x, rep = solve_mixed(box["asm"], rtol=1e-12, storage=storage, coarse="rigid",
                         precond_bj_chunk=150)
    err = np.linalg.norm(x.cpu().numpy() - box["x_lu"]) / np.linalg.norm(box["x_lu"])
    print(f"{rep}; err {err:.1e}")
    assert rep.converged and err < 1e-9


def test_mixed_operator_image(box):
    """The coarse image of the mixed operator (int8 far field + FP64 near
    field) equals that operator applied to every column of W, to the FP32
    accumulation of its matvec."""
    import torch
    from mbem_amd import _lib
    from mbem_amd.la.coarse import CoarseSpace
    from mbem_amd.la.mixed import MixedOperator
    asm, M = box["asm"], box["M"]
    pts, kind, comp = _geometry(asm)
    W = oc.coarse_basis(M.chunk_dofs, pts, kind, comp, "rigid")
    op = MixedOperator(asm, storage="int8")
    cs = CoarseSpace(asm, M.chunk_dofs, basis="rigid", op=op)
    assert cs.source.startswith("int8")
    Ct = cs._Ct.cpu().numpy()                    # before finish() drops it
    dev = asm.A_dev.device
    ref = np.stack([op.matvec(torch.as_tensor(np.ascontiguousarray(W[:, c]), device=dev)).cpu().numpy()
                    for c in range(W.shape[1])])
    err = np.abs(Ct - ref).max() / np.abs(ref).max()
    print(f"int8 image vs the mixed matvec: {err:.1e}")
    assert err < 1e-5
    cs.finish()
    assert cs.Kp == cs.K
    cs.check()
