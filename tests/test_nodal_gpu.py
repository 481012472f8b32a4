"""The Lagrange P1/P2 nodal kernels on the GPU (csrc/nodal.hip through
mbem_amd.kernels.tri_nodal) against the pinned nodal oracle
(oracle/tri_nodal.py, tests/test_nodal_oracle.py) on the same inputs: a
warped 32-triangle patch with field points on it (0.02 and 0.3 of a size
above centroids, at a vertex), beside it, and in both far-field Gauss zones
(10 L < D <= 40 L, D > 40 L), varying eps per element, every driver:

* U and T matrices (coefficients applied) and basis stacks;
* displacement contractions of force and slip densities;
* force -> stress, slip -> total stress and eigenstress contractions;
* traction rows of every kind at orders 0, 1 and 2, with the element-major
  ids and with a shared-node table (and through basis.assemble_traction_matrix).

Bar: 1e-10 of each result's largest entry (the kernels round as the oracle's
restatement does; measured differences are printed).
"""
import numpy as np
import pytest

from oracle import tri_nodal as on

pytestmark = pytest.mark.gpu

MU, LAM = 30.0, 45.0
BAR = 1e-10


def _patch():
    m = 4
    xs = np.linspace(0.0, 3.0, m + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    V = np.stack([X, Y, 0.15 * np.sin(X) * np.cos(Y)], axis=-1)
    tris = []
    for i in range(m):
        for j in range(m):
            a, b, c, d = V[i, j], V[i + 1, j], V[i + 1, j + 1], V[i, j + 1]
            tris += [np.array([a, b, c]), np.array([a, c, d])]
    tv = np.array(tris)
    nrm = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rng = np.random.default_rng(7)
    eps = rng.uniform(0.2, 0.45, size=tv.shape[0])
    cen = tv.mean(axis=1)
    x = np.concatenate([cen[::3] + 0.02 * nrm[::3], cen[1::5] + 0.3 * nrm[1::5], tv[4, 1:2],
                        rng.uniform(-1.0, 4.0, (5, 3)), cen[:2] + np.array([25.0, 3.0, 1.0]),
                        cen[2:4] + np.array([90.0, -40.0, 8.0])])
    return tv, nrm, eps, x, rng


def _check(name, got, ref):
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"{name}: {err:.1e}", flush=True)
    assert got.shape == ref.shape and err < BAR, (name, err)


def _coeffs():
    from mbem_amd.kernels.basis import t_coeffs, u_coeffs
    return u_coeffs(MU, LAM), t_coeffs(MU, LAM)


@pytest.mark.parametrize("order", [1, 2])
def test_matrices_and_stacks_match_oracle(order):
    from mbem_amd.kernels import tri_nodal as tn
    tv, nrm, eps, x, _ = _patch()
    g, c = _coeffs()
    print(f"[nodal order {order}] matrices: {x.shape[0]} points x {tv.shape[0]} elements",
          flush=True)
    _check("u_matrix_direct", tn.u_matrix_direct(x, tv, eps, *g, order),
           on.u_matrix_direct(x, tv, eps, *g, order))
    _check("t_matrix_direct", tn.t_matrix_direct(x, tv, nrm, eps, *c, order),
           on.t_matrix_direct(x, tv, nrm, eps, *c, order))
    _check("u_basis_matrices", tn.u_basis_matrices(x, tv, eps, order),
           on.u_basis_matrices(x, tv, eps, order))
    _check("t_basis_matrices", tn.t_basis_matrices(x, tv, nrm, eps, order),
           on.t_basis_matrices(x, tv, nrm, eps, order))


@pytest.mark.parametrize("order", [1, 2])
def test_contractions_match_oracle(order):
    from mbem_amd.kernels import tri_nodal as tn
    tv, nrm, eps, x, rng = _patch()
    g, c = _coeffs()
    K = tn.n_nodes(order)
    dens = rng.normal(size=(K * tv.shape[0], 3))
    dens[3 * K:4 * K] = 0.0                 # a zero block is skipped
    print(f"[nodal order {order}] contractions", flush=True)
    _check("u_disp_contract", tn.u_disp_contract(x, tv, eps, dens, *g, order),
           on.u_disp_contract(x, tv, eps, dens, *g, order))
    _check("t_disp_contract", tn.t_disp_contract(x, tv, nrm, eps, dens, *c, order),
           on.t_disp_contract(x, tv, nrm, eps, dens, *c, order))
    _check("kelvin_stress_contract", tn.kelvin_stress_contract(x, tv, eps, dens, MU, LAM, order),
           on.kelvin_stress_contract(x, tv, eps, dens, MU, LAM, order))
    _check("dd_stress_contract", tn.dd_stress_contract(x, tv, nrm, eps, dens, MU, LAM, order),
           on.dd_stress_contract(x, tv, nrm, eps, dens, MU, LAM, order))
    _check("eigenstress_contract", tn.eigenstress_contract(x, tv, nrm, eps, dens, MU, LAM, order),
           on.eigenstress_contract(x, tv, nrm, eps, dens, MU, LAM, order))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_traction_rows_match_oracle(order):
    from mbem_amd.kernels import tri_nodal as tn
    tv, nrm, eps, x, rng = _patch()
    K = tn.n_nodes(order)
    ns = tv.shape[0]
    nf = rng.normal(size=x.shape)
    nf /= np.linalg.norm(nf, axis=1, keepdims=True)
    ids = np.arange(K * ns).reshape(ns, K)
    shared = ids.copy()
    shared[5] = shared[4]
    shared[9, 0] = shared[2, 0]
    for which in ("slip", "force", "eigen", "total"):
        print(f"[nodal order {order}] traction rows {which}", flush=True)
        _check(f"traction {which}", tn.traction_matrix(x, nf, tv, nrm, eps, ids, K * ns, MU, LAM,
                                                       order, which),
               on.traction_matrix(x, nf, tv, nrm, eps, ids, K * ns, MU, LAM, order, which))
    print(f"[nodal order {order}] traction rows, shared nodes", flush=True)
    _check("traction shared", tn.traction_matrix(x, nf, tv, nrm, eps, shared, K * ns, MU, LAM,
                                                 order, "slip"),
           on.traction_matrix(x, nf, tv, nrm, eps, shared, K * ns, MU, LAM, order, "slip"))


def test_assemble_traction_matrix_through_basis():
    """basis.assemble_traction_matrix (basis.py:406-427) on a TriMesh: the
    element-major layout by default, the rows of an elastic material."""
    from mbem_amd.kernels.basis import assemble_traction_matrix
    from mbem_amd.mesh import ElasticMaterial, TriMesh
    tv, nrm, eps, x, rng = _patch()
    verts = tv.reshape(-1, 3)
    mesh = TriMesh(verts, np.arange(verts.shape[0]).reshape(-1, 3))
    mat = ElasticMaterial(mu=MU, lam=LAM)
    nf = np.tile([0.0, 0.0, 1.0], (x.shape[0], 1))
    got = assemble_traction_matrix(x, nf, mesh, mat, eps, order=1, kind="slip")
    ids = np.arange(3 * tv.shape[0]).reshape(-1, 3)
    _check("assemble_traction_matrix", got,
           on.traction_matrix(x, nf, tv, nrm, eps, ids, ids.size, mat.mu, mat.lam, 1, "slip"))
