"""kernels.tri_nodal, the nodal entry points with a trailing order
(tri_nodal.py:51-56, 1257-1404): order bookkeeping on the host, and on the
GPU the order-0 wrappers equal to kernels.tri_kernels on the same inputs
(the reference routes order 0 to tri_kernels the same way)."""
import numpy as np
import pytest

from _geom import normals_of, random_triangles


def test_orders_on_the_host():
    """kernels.tri_nodal (tri_nodal.py:51-56, 1239-1404): n_nodes 1 / 3 / 6,
    a malformed order is a ValueError, and before any device call a needle
    (height / L < 1e-3) is refused at orders 1 and 2, and traction_matrix
    refuses an ids table of the wrong shape or range."""
    from mbem_amd.kernels import tri_nodal as tn
    assert [tn.n_nodes(o) for o in (0, 1, 2)] == [1, 3, 6]
    with pytest.raises(ValueError):
        tn.n_nodes(3)
    needle = np.array([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1e-4, 0.0]]])
    for order in (1, 2):
        with pytest.raises(ValueError, match="height / L"):
            tn.u_matrix_direct(np.zeros((1, 3)), needle, np.ones(1), 1.0, 1.0, 1.0, order)
    tv = np.zeros((1, 3, 3))
    with pytest.raises(ValueError):
        tn.dd_stress_contract(np.zeros((1, 3)), tv, tv[:, 0], np.ones(1), np.ones((1, 3)),
                              30.0, 30.0, 5)
    with pytest.raises(ValueError, match="ids shape"):
        tn.traction_matrix(np.zeros((1, 3)), np.zeros((1, 3)), tv, tv[:, 0], np.ones(1),
                           np.zeros((1, 3), dtype=np.int64), 3, 30.0, 30.0, 0)
    with pytest.raises(ValueError, match="ids must index"):
        tn.traction_matrix(np.zeros((1, 3)), np.zeros((1, 3)), tv, tv[:, 0], np.ones(1),
                           np.full((1, 1), 4, dtype=np.int64), 2, 30.0, 30.0, 0)


def test_nodal_spec_matches_the_oracle_spec():
    """The device spec's degrees, series coefficients and Gauss tables equal
    the oracle's restatement of tri_nodal.py:77-221 for every kernel and
    order (both restate the same rules; the oracle's spec is what its pinned
    kernels run on)."""
    from mbem_amd.kernels import nodal as nd
    from oracle import tri_nodal as on
    for order in (0, 1, 2):
        for flags in (nd.KT, nd.KU, nd.KH, nd.KS, nd.KE):
            sp = nd.Spec(order, flags)
            isp, fsp, lat, binom, q_lam, q_w, q_N = on.spec(order, flags)
            assert list(sp.ispec[:3]) == [order, sp.K, isp[2]]
            assert list(sp.ispec[5:10]) == list(isp[3:8])
            assert list(sp.ispec[10:14]) == list(isp[8:12])
            assert list(sp.ispec[14:18]) == list(isp[12:16])
            assert np.array_equal(sp.fspec, fsp)
            nt, nq, K = int(sp.ispec[3]), int(sp.ispec[4]), sp.K
            t = sp._host_tables
            parts = np.split(t, np.cumsum([4 * nt, 3 * K, 3 * nq, nq]))
            assert np.array_equal(parts[0].reshape(4, nt), binom)
            assert np.array_equal(parts[1].reshape(K, 3), lat.astype(float))
            assert np.array_equal(parts[2].reshape(nq, 3), q_lam)
            assert np.array_equal(parts[3], q_w)
            assert np.array_equal(parts[4].reshape(nq, K), q_N)


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
