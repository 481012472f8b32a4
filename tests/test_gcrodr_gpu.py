"""GPU parity of the recycled solves (la/solver.py:248-500; the sequence
options of hmat.py:605-664) on the 2,048-triangle box, against the oracle's
GCRO-DR (oracle/dense.py gcrodr, itself pinned in test_gcrodr_oracle.py).

Bars: over a sequence of three systems (A + sigma_i diag(A), sigma_i = 0,
0.05, 0.1, the model's right-hand side scaled; measured with the oracle:
56, 44, 38 iterations recycled against 60, 56, 54 for FGMRES(30)),
the patch block Gauss-Seidel preconditioner, restart 30 so the deflated
restarts and the carried space are exercised) the device solve converges to
true residual 1e-10 in each, with iteration counts within 2 of the oracle's
and solutions within 1e-8 (relative 2-norm) of numpy's; recycling never
costs iterations against the same sequence without it (the reference's own
gate, defaults.py:100-109).  Along a ``rebuild_for_materials`` chain,
``recycle=True`` + ``x0="previous"`` solve each step to the LU solution,
in no more iterations than a fresh solve inside the reuse window, and a
material step past log 2 drops the space (no setup matvecs on that solve).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def box():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    from oracle import dense as od
    model, eps, _ = box_fault_model(16)
    system = generate_system(model)
    asm = DenseBackend().assemble(system, eps)
    A, b, _ = od.assemble(system, eps)
    return dict(model=model, eps=eps, system=system, asm=asm, A=A, b=b)


def test_sequence_matches_oracle(box):
    from mbem_amd.la.preconditioner import BlockGaussSeidel
    from mbem_amd.la.solver import RecycleSpace, fgmres
    from oracle import dense as od
    asm, A = box["asm"], box["A"]
    n = A.shape[0]
    dev = asm.A_dev.device
    pts = {id(p): p.collocation_points() for p in box["model"].regions[0].patches}
    M_or = od.BlockJacobiGS(A, asm.layout, lambda p: pts[id(p)], max_dense=0, bj_chunk=1500)
    M = BlockGaussSeidel(asm, max_dense=0, bj_chunk=1500, grouping="patch")
    dA = np.diag(np.diag(A))
    rs = RecycleSpace(dim=10)
    Y_or = None
    tot_rec, tot_plain = 0, 0
    for i, sigma in enumerate((0.0, 0.05, 0.1)):
        As = A + sigma * dA
        Ad = torch.as_tensor(As, device=dev)
        b = box["b"] * (1.0 + 0.5 * i)
        bd = torch.as_tensor(b, device=dev)
        x, rep = fgmres(Ad, bd, M=M, rtol=1e-10, restart=30, maxiter=600, recycle=rs)
        xo, io = od.gcrodr(lambda v: As @ v, b, M=M_or, Y=Y_or, dim=10, rtol=1e-10, restart=30)
        Y_or = io["Y"]
        _, plain = fgmres(Ad, bd, M=M, rtol=1e-10, restart=30, maxiter=600)
        ref = np.linalg.solve(As, b)
        e = np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"system {i}: GPU {rep.iterations} it (recycle dim {rep.recycle_dim}, setup "
              f"{rep.setup_matvecs}), oracle {io['iterations']}, without recycling "
              f"{plain.iterations}; err vs numpy {e:.1e}", flush=True)
        assert rep.converged and rep.true_relres < 1e-10
        assert abs(rep.iterations - io["iterations"]) <= 2
        assert e < 1e-8
        tot_rec += rep.iterations
        tot_plain += plain.iterations
        del Ad
    assert tot_rec <= tot_plain


def test_lineage_warm_start_recycle_and_reset(box):
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.mesh import ElasticMaterial
    asm0 = box["asm"]
    region = box["model"].regions[0].name
    chain = asm0
    iters = []
    for mu in (30.0, 33.0, 36.0, 120.0):
        chain = chain.rebuild_for_materials({region: ElasticMaterial(mu=mu, lam=30.0)})
        x = chain.solve(rtol=1e-11, restart=40, recycle=True, x0="previous",
                        precond_bj_chunk=1500, return_device=True)
        rep = chain.report
        x_lu = chain.solve(method="lu", return_device=True)
        e = float(torch.linalg.vector_norm(x - x_lu) / torch.linalg.vector_norm(x_lu))
        fresh = AssembledDense(box["system"], box["eps"], _materials=chain.materials)
        fresh.solve(rtol=1e-11, restart=40, precond_bj_chunk=1500)
        print(f"mu {mu}: recycled + warm {rep.iterations} it (setup {rep.setup_matvecs}, dim "
              f"{rep.recycle_dim}), fresh {fresh.report.iterations}; vs LU {e:.1e}", flush=True)
        assert rep.converged and e < 1e-9
        if mu < 100.0:               # inside the reuse window
            assert rep.iterations <= fresh.report.iterations
        iters.append((mu, rep.setup_matvecs))
        chain.release()
        del fresh
    # 36 -> 120 is a step of log(10/3) > log 2: the space is dropped first
    assert iters[1][1] > 0 and iters[2][1] > 0 and iters[3][1] == 0
