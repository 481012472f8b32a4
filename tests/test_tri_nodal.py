"""kernels.tri_nodal, the nodal entry points with a trailing order
(tri_nodal.py:51-56, 1257-1404): order bookkeeping on the host, and on the
GPU the order-0 wrappers equal to kernels.tri_kernels on the same inputs
(the reference routes order 0 to tri_kernels the same way)."""
import numpy as np
import pytest

from _geom import normals_of, random_triangles


def test_orders_on_the_host():
    """kernels.tri_nodal (tri_nodal.py:51-56, 1257-1404): n_nodes 1 / 3 / 6,
    a malformed order is a ValueError, orders 1 and 2 and traction_matrix are
    refused before any device call."""
    from mbem_amd.kernels import tri_nodal as tn
    assert [tn.n_nodes(o) for o in (0, 1, 2)] == [1, 3, 6]
    with pytest.raises(ValueError):
        tn.n_nodes(3)
    tv = np.zeros((1, 3, 3))
    with pytest.raises(NotImplementedError):
        tn.u_matrix_direct(np.zeros((1, 3)), tv, np.ones(1), 1.0, 1.0, 1.0, 1)
    with pytest.raises(ValueError):
        tn.dd_stress_contract(np.zeros((1, 3)), tv, tv[:, 0], np.ones(1), np.ones((1, 3)),
                              30.0, 30.0, 5)
    with pytest.raises(NotImplementedError):
        tn.traction_matrix(np.zeros((1, 3)), np.zeros((1, 3)), tv, tv[:, 0], np.ones(1),
                           np.zeros((1, 1), dtype=np.int64), 1, 30.0, 30.0, 0)


@pytest.mark.gpu
def test_order0_routes_to_the_pair_kernels():
    from mbem_amd.kernels import tri_kernels as tk
    from mbem_amd.kernels import tri_nodal as tn
    rng = np.random.default_rng(31)
    tv = random_triangles(rng, 9)
    nrm = normals_of(tv)
    eps = rng.uniform(0.1, 0.5, size=9)
    pts = rng.uniform(-2.0, 2.0, size=(11, 3))
    dens = rng.normal(size=(9, 3))
    g, c = (1.0, 0.3, 0.7), (0.2, -0.1, 0.4, 0.9, -0.3, 0.5)
    pairs = [
        (tn.u_matrix_direct(pts, tv, eps, *g, 0), tk.u_matrix_direct(pts, tv, eps, *g)),
        (tn.t_matrix_direct(pts, tv, nrm, eps, *c, 0), tk.t_matrix_direct(pts, tv, nrm, eps, *c)),
        (tn.u_basis_matrices(pts, tv, eps, 0), tk.u_basis_matrices(pts, tv, eps)),
        (tn.t_basis_matrices_serial(pts, tv, nrm, eps, 0),
         tk.t_basis_matrices_serial(pts, tv, nrm, eps)),
        (tn.t_disp_contract(pts, tv, nrm, eps, dens, *c, 0),
         tk.t_disp_contract(pts, tv, nrm, eps, dens, *c)),
        (tn.dd_stress_contract(pts, tv, nrm, eps, dens, 30.0, 30.0, 0),
         tk.dd_stress_contract(pts, tv, nrm, eps, dens, 30.0, 30.0)),
        (tn.eigenstress_contract(pts, tv, nrm, eps, dens, 30.0, 30.0, 0),
         tk.eigenstress_contract(pts, tv, nrm, eps, dens, 30.0, 30.0)),
    ]
    for a, b in pairs:
        assert a.shape == b.shape and np.array_equal(a, b)
