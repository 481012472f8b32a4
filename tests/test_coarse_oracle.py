"""The two-level block-Jacobi oracle (oracle/coarse.py) pinned on the CPU,
and the product's host-side coarse basis checked against it.

* deflation identity: for r = A w with w in span W the preconditioner
  returns w itself (to rounding);
* the basis: per (chunk, kind) group the constants are the group's
  indicator times e_b, and constants + rotations hold every rigid motion
  c + omega x p of the group's points exactly;
* FGMRES (the oracle's restatement of the reference loop) with the
  two-level preconditioner reaches the LU solution, in fewer iterations
  than block-Jacobi alone;
* mbem_amd.la.coarse builds the same W as the oracle (its permuted,
  compact form expanded).
"""
import numpy as np
import pytest

from oracle import coarse as oc
from oracle import dense as od


@pytest.fixture(scope="module")
def small():
    from mbem_amd.cases import box_fault_model
    from mbem_amd.la.cluster import build_cluster_tree, leaves
    from mbem_amd.model import generate_system
    model, eps, _ = box_fault_model(8, fault_cells=(6, 10))
    system = generate_system(model)
    A, b, _ = od.assemble(system, eps)
    layout = system.layout
    n = layout.n_unknowns
    pts = np.empty((n, 3))
    kind = np.empty(n, np.int64)
    for s in layout.slots:
        pts[s.offset:s.stop] = np.repeat(s.patch.collocation_points(), 3, axis=0)
        kind[s.offset:s.stop] = 0 if s.kind == "u" else 1
    comp = np.tile(np.arange(3), n // 3)
    tri_pts = pts[::3]
    chunks = [np.asarray(c) for c in leaves(build_cluster_tree(tri_pts, 50))]
    chunk_dofs = [(3 * c[:, None] + np.arange(3)[None, :]).ravel() for c in chunks]
    return dict(A=A, b=b, pts=pts, kind=kind, comp=comp, chunk_dofs=chunk_dofs, n=n)


@pytest.mark.parametrize("basis", ["const", "rigid", "linear"])
def test_deflation_identity(small, basis):
    A = small["A"]
    W = oc.coarse_basis(small["chunk_dofs"], small["pts"], small["kind"], small["comp"], basis)
    P = oc.two_level(A, W, oc.block_jacobi(A, small["chunk_dofs"]))
    rng = np.random.default_rng(0)
    w = W @ rng.standard_normal(W.shape[1])
    z = P(A @ w)
    assert np.linalg.norm(z - w) < 1e-9 * np.linalg.norm(w)


def test_basis_holds_rigid_motions(small):
    W = oc.coarse_basis(small["chunk_dofs"], small["pts"], small["kind"], small["comp"], "rigid")
    pts, kind, comp = small["pts"], small["kind"], small["comp"]
    rng = np.random.default_rng(1)
    col = 0
    for dofs in small["chunk_dofs"]:
        for k in (0, 1):
            d = dofs[kind[dofs] == k]
            if d.size == 0:
                continue
            Wg = W[d, col:col + 6]
            for b in range(3):
                assert np.array_equal(Wg[:, b], (comp[d] == b).astype(float))
            c, om = rng.standard_normal(3), rng.standard_normal(3)
            u = (c[None, :] + np.cross(om[None, :], pts[d] - pts[d].mean(0)))[np.arange(d.size),
                                                                            comp[d]]
            coef, *_ = np.linalg.lstsq(Wg, u, rcond=None)
            assert np.linalg.norm(Wg @ coef - u) < 1e-12 * np.linalg.norm(u)
            assert not np.any(W[np.setdiff1d(np.arange(small["n"]), d), col:col + 6])
            col += 6
    assert col == W.shape[1]


def test_two_level_fgmres_reaches_lu(small):
    A, b = small["A"], small["b"]
    M = oc.block_jacobi(A, small["chunk_dofs"])
    W = oc.coarse_basis(small["chunk_dofs"], small["pts"], small["kind"], small["comp"], "rigid")
    P = oc.two_level(A, W, M)
    x_lu = od.lu_solve(A, b)
    _, rep_bj = od.fgmres(lambda v: A @ v, b, M=M, rtol=1e-12, restart=300, maxiter=300)
    x, rep = od.fgmres(lambda v: A @ v, b, M=P, rtol=1e-12, restart=300, maxiter=300)
    print(f"block-Jacobi {rep_bj['iterations']}, two-level {rep['iterations']}")
    assert rep["converged"]
    assert np.linalg.norm(x - x_lu) < 1e-9 * np.linalg.norm(x_lu)
    assert rep["iterations"] < rep_bj["iterations"]


@pytest.mark.parametrize("basis", ["const", "rigid", "linear"])
def test_product_basis_equals_oracle(small, basis):
    from mbem_amd.la.coarse import coarse_groups
    perm, gptr, Wp = coarse_groups(small["chunk_dofs"], small["pts"], small["kind"],
                                   small["comp"], basis)
    q = Wp.shape[1]
    W = np.zeros((small["n"], (gptr.size - 1) * q))
    for g in range(gptr.size - 1):
        t = slice(gptr[g], gptr[g + 1])
        W[perm[t], g * q:(g + 1) * q] = Wp[t]
    Wo = oc.coarse_basis(small["chunk_dofs"], small["pts"], small["kind"], small["comp"], basis)
    assert W.shape == Wo.shape
    assert np.abs(W - Wo).max() < 1e-14
