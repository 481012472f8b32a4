"""The nodal oracle (oracle/tri_nodal_ref.c, oracle/tri_nodal.py) pinned on
the CPU before it is trusted.

* Gauss product quadrature of the POINTWISE mollified Kelvin kernel and its
  symbolic derivatives (tests/_quad.py) weighted by the textbook Lagrange
  basis of the element (P1: lam_k; P2: lam_i (2 lam_i - 1), 4 lam_i lam_j),
  written without the node-lattice product formula the kernels use: every
  node's U, T, force->stress, slip->stress and eigenstress columns at points
  on, above, beside, outside and far from the element (both far-field Gauss
  rules), to 1e-11 relative, the quadrature converged first;
* the partition of unity sum_k N_k = 1: the node columns of every driver
  summed over an element equal the P0 restatement (oracle/tri_kernels_ref.c,
  itself pinned in test_oracle_pinning.py) to 1e-11 of the largest entry;
* the traction rows at order 0 (which run the nodal machinery) equal the P0
  stress contractions dotted with the field normals, for every kind; at
  orders 1 and 2 equal the nodal stress contractions; a shared-node id table
  sums the columns of coincident nodes;
* needles (height / L < 1e-3) are refused;
* a planted mistake (the sign of one barycentric gradient in the shape
  coefficients) is caught by the quadrature comparison.
"""
import os
import subprocess

import numpy as np
import pytest

import _quad
from oracle import tri_kernels as ok
from oracle import tri_nodal as on

TOL = 1e-11
MU, LAM, EPS = 1.3, 2.1, 0.4


def _tri():
    return np.array([[0.0, 0.0, 0.0], [2.0, 0.3, 0.1], [0.4, 1.7, -0.2]])


def _normal(tv):
    c = np.cross(tv[1] - tv[0], tv[2] - tv[0])
    return c / np.linalg.norm(c)


def _coeffs(mu, lam):
    from mbem_amd.kernels.basis import t_coeffs, u_coeffs
    return u_coeffs(mu, lam), t_coeffs(mu, lam)


def _points(tv, n):
    c = tv.mean(axis=0)
    # on the element, above it, by a vertex, outside, and in both far-field
    # Gauss zones (10 L < D <= 40 L and beyond)
    return [c, c + 0.3 * n, tv[0] + 0.05 * n, tv[1] + np.array([0.5, 0.0, 0.0]),
            np.array([3.0, 2.0, 1.0]), c + np.array([30.0, 0.0, 0.0]),
            c + np.array([100.0, 20.0, 5.0])]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("order", [1, 2])
def test_displacement_nodes_match_weighted_quadrature(order):
    tv = _tri()
    n = _normal(tv)
    g, c = _coeffs(MU, LAM)
    K = on.n_nodes(order)
    for obs in _points(tv, n):
        U = on.u_matrix_direct(obs[None], tv[None], [EPS], *g, order)
        T = on.t_matrix_direct(obs[None], tv[None], n[None], [EPS], *c, order)
        for k in range(K):
            w = _quad.lagrange_weight(tv, order, k)
            qa, q = _quad.pair("G", tv, obs, EPS, MU, LAM, depth=2, n=16, weight=w)
            assert _rel(qa, q) < 1e-12
            # a P2 vertex function integrates to almost nothing far away:
            # measure against the element's whole P0 block there
            scale = max(np.linalg.norm(q), 1e-3 * np.linalg.norm(U.reshape(3, K, 3).sum(1)))
            assert np.linalg.norm(U[:, 3 * k:3 * k + 3] - q) < TOL * scale, (order, k, obs)
            qa, q = _quad.pair("T", tv, obs, EPS, MU, LAM, depth=2, n=16, weight=w)
            assert _rel(qa, q) < 1e-12
            scale = max(np.linalg.norm(q), 1e-3 * np.linalg.norm(T.reshape(3, K, 3).sum(1)))
            assert np.linalg.norm(T[:, 3 * k:3 * k + 3] - q) < TOL * scale, (order, k, obs)


@pytest.mark.parametrize("order", [1, 2])
def test_stress_nodes_match_weighted_quadrature(order):
    tv = _tri()
    n = _normal(tv)
    K = on.n_nodes(order)
    d = np.array([0.3, -0.7, 0.5])
    pts = _points(tv, n)
    for obs in (pts[1], pts[2], pts[4], pts[5]):
        for k in range(K):
            dens = np.zeros((K, 3))
            dens[k] = d
            w = _quad.lagrange_weight(tv, order, k)
            for kind, fn in (("S", lambda: on.kelvin_stress_contract(obs[None], tv[None], [EPS],
                                                                      dens, MU, LAM, order)),
                             ("H", lambda: on.dd_stress_contract(obs[None], tv[None], n[None],
                                                                  [EPS], dens, MU, LAM, order)),
                             ("E", lambda: on.eigenstress_contract(obs[None], tv[None], n[None],
                                                                    [EPS], dens, MU, LAM, order))):
                qa, q = _quad.pair(kind, tv, obs, EPS, MU, LAM, depth=2, n=16, weight=w)
                ref = q @ d
                got = fn()[0]
                assert _rel(qa @ d, ref) < 1e-12
                assert _rel(got, ref) < TOL, (order, k, kind, obs)


def _mesh(rng, m=4):
    """m x m structured patch, slightly warped, as 2 m^2 triangles."""
    xs = np.linspace(0.0, 3.0, m + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    Z = 0.15 * np.sin(X) * np.cos(Y)
    V = np.stack([X, Y, Z], axis=-1)
    tris = []
    for i in range(m):
        for j in range(m):
            a, b, c, dd = V[i, j], V[i + 1, j], V[i + 1, j + 1], V[i, j + 1]
            tris += [np.array([a, b, c]), np.array([a, c, dd])]
    tv = np.array(tris)
    nrm = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return tv, nrm


@pytest.mark.parametrize("order", [1, 2])
def test_partition_of_unity_equals_p0(order):
    rng = np.random.default_rng(3)
    tv, nrm = _mesh(rng)
    ns = tv.shape[0]
    K = on.n_nodes(order)
    eps = np.full(ns, 0.3)
    cen = tv.mean(axis=1)
    x = np.concatenate([cen[::5] + 0.02 * nrm[::5], rng.uniform(-1, 4, (6, 3)),
                        cen[:3] + np.array([40.0, 0.0, 0.0])])
    g, c = _coeffs(MU, LAM)
    sumk = lambda A: A.reshape(A.shape[:-1] + (ns, K, 3)).sum(-2).reshape(A.shape[:-1] + (3 * ns,))  # noqa: E731

    # 1e-11 of the largest entry: beyond 10 L the nodal path integrates by
    # Gauss, while the P0 closed form's cancelling terms lose digits there
    # (3e-7 of a 3.5e-6 entry at 38 L in the T stack; the nodal columns
    # match weighted quadrature at such points to 1e-11 above)
    def close(a, b):
        return np.abs(a - b).max() < 1e-11 * np.abs(b).max()
    assert close(sumk(on.u_matrix_direct(x, tv, eps, *g, order)), ok.u_matrix_direct(x, tv, eps, *g))
    assert close(sumk(on.t_matrix_direct(x, tv, nrm, eps, *c, order)),
                 ok.t_matrix_direct(x, tv, nrm, eps, *c))
    assert close(sumk(on.u_basis_matrices(x, tv, eps, order)), ok.u_basis_matrices(x, tv, eps))
    assert close(sumk(on.t_basis_matrices(x, tv, nrm, eps, order)),
                 ok.t_basis_matrices(x, tv, nrm, eps))
    # a density equal at every node of an element is the P0 density
    d0 = rng.normal(size=(ns, 3))
    dn = np.repeat(d0, K, axis=0)
    assert close(on.u_disp_contract(x, tv, eps, dn, *g, order), ok.u_disp_contract(x, tv, eps, d0, *g))
    assert close(on.t_disp_contract(x, tv, nrm, eps, dn, *c, order),
                 ok.t_disp_contract(x, tv, nrm, eps, d0, *c))
    assert close(on.kelvin_stress_contract(x, tv, eps, dn, MU, LAM, order),
                 ok.kelvin_stress_contract(x, tv, eps, d0, MU, LAM))
    assert close(on.dd_stress_contract(x, tv, nrm, eps, dn, MU, LAM, order),
                 ok.dd_stress_contract(x, tv, nrm, eps, d0, MU, LAM))
    assert close(on.eigenstress_contract(x, tv, nrm, eps, dn, MU, LAM, order),
                 ok.eigenstress_contract(x, tv, nrm, eps, d0, MU, LAM))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_traction_rows_equal_stress_contractions(order):
    rng = np.random.default_rng(4)
    tv, nrm = _mesh(rng, 3)
    ns = tv.shape[0]
    K = on.n_nodes(order)
    eps = np.full(ns, 0.35)
    x = np.concatenate([tv.mean(axis=1)[::4] + 0.05 * nrm[::4], rng.uniform(-1, 4, (3, 3))])
    nf = rng.normal(size=x.shape)
    nf /= np.linalg.norm(nf, axis=1, keepdims=True)
    ids = np.arange(K * ns).reshape(ns, K)
    dens = rng.normal(size=(K * ns, 3))
    if order == 0:
        st = {"force": ok.kelvin_stress_contract(x, tv, eps, dens, MU, LAM),
              "eigen": ok.eigenstress_contract(x, tv, nrm, eps, dens, MU, LAM),
              "total": ok.dd_stress_contract(x, tv, nrm, eps, dens, MU, LAM)}
    else:
        st = {"force": on.kelvin_stress_contract(x, tv, eps, dens, MU, LAM, order),
              "eigen": on.eigenstress_contract(x, tv, nrm, eps, dens, MU, LAM, order),
              "total": on.dd_stress_contract(x, tv, nrm, eps, dens, MU, LAM, order)}
    st["slip"] = st["total"] - st["eigen"]
    for which, sig in st.items():
        Tm = on.traction_matrix(x, nf, tv, nrm, eps, ids, K * ns, MU, LAM, order, which)
        t = (Tm @ dens.ravel()).reshape(-1, 3)
        ref = np.einsum("fij,fj->fi", sig, nf)
        assert np.abs(t - ref).max() < 1e-12 * np.abs(ref).max(), (order, which)
    # shared nodes: two elements' nodes mapped onto one column add up
    ids2 = ids.copy()
    ids2[1] = ids2[0]
    T1 = on.traction_matrix(x, nf, tv, nrm, eps, ids, K * ns, MU, LAM, order, "slip")
    T2 = on.traction_matrix(x, nf, tv, nrm, eps, ids2, K * ns, MU, LAM, order, "slip")
    cols = lambda k: slice(3 * k, 3 * k + 3)  # noqa: E731
    for k in range(K):
        assert np.allclose(T2[:, cols(ids[0, k])], T1[:, cols(ids[0, k])] + T1[:, cols(ids[1, k])],
                           rtol=0, atol=1e-14 * np.abs(T1).max())
        assert not np.any(T2[:, cols(ids[1, k])])


def test_needles_are_refused():
    tv = np.array([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1e-4, 0.0]]])
    with pytest.raises(ValueError, match="height / L"):
        on.u_matrix_direct(np.zeros((1, 3)), tv, [0.1], 1.0, 1.0, 1.0, 1)


def test_planted_error_is_caught(tmp_path):
    """Rebuild the nodal oracle with the sign of one barycentric gradient
    flipped in the shape coefficients and show the quadrature check fails."""
    here = os.path.dirname(on.__file__)
    text = open(os.path.join(here, "tri_nodal_ref.c")).read()
    old = "fac[0][1] = (order * geo[20 + 3 * i]) / (j + 1);"
    assert old in text
    planted = text.replace(old, "fac[0][1] = -(order * geo[20 + 3 * i]) / (j + 1);")
    csrc = tmp_path / "planted_nodal.c"
    csrc.write_text(planted)
    so = tmp_path / "planted_nodal.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    f"-I{here}", "-o", str(so), str(csrc), "-lm"], check=True)
    good = on._lib
    try:
        on._lib = None
        path = on._LIB_PATH
        on._LIB_PATH = str(so)
        tv = _tri()
        n = _normal(tv)
        obs = tv.mean(axis=0) + 0.3 * n
        g, _ = _coeffs(MU, LAM)
        U = on.u_matrix_direct(obs[None], tv[None], [EPS], *g, 1)
    finally:
        on._lib = good
        on._LIB_PATH = path
    _, q = _quad.pair("G", tv, obs, EPS, MU, LAM, depth=2, n=16,
                      weight=_quad.lagrange_weight(tv, 1, 0))
    assert _rel(U[:, 0:3], q) > 1e-6
    assert _rel(on.u_matrix_direct(obs[None], tv[None], [EPS], *g, 1)[:, 0:3], q) < TOL
