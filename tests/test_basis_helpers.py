"""Host helpers of kernels.basis (basis.py:91-301) pinned by what they must
satisfy: Lagrange shape functions interpolate (N_k(node_l) = delta_kl) and
sum to one everywhere, for orders 0-2; nodes shrink toward the centroid on
the segment the formula states; condense_columns adds element-major columns
into shared node columns exactly (checked against an explicit scatter
matrix), accumulates into ``out``, and refuses mismatched shapes and ids;
MeshArrays computes each array once."""
import numpy as np
import pytest

from mbem_amd.kernels import basis as kb


@pytest.mark.parametrize("order", [0, 1, 2])
def test_lagrange_interpolation_and_partition_of_unity(order):
    nodes = kb.lagrange_nodes(order)
    N = kb.lagrange_shape(order, nodes)
    assert np.allclose(N, np.eye(nodes.shape[0]), atol=1e-15)
    rng = np.random.default_rng(order)
    lam = rng.dirichlet(np.ones(3), size=50)
    assert np.allclose(kb.lagrange_shape(order, lam).sum(axis=1), 1.0, atol=1e-14)
    t = 0.25
    sh = kb.lagrange_nodes(order, shrink=t)
    assert np.allclose(sh, (1 - t) * nodes + t / 3.0) and np.allclose(sh.sum(axis=1), 1.0)
    with pytest.raises(ValueError):
        kb.lagrange_nodes(3)


def test_condense_columns_against_a_scatter_matrix():
    rng = np.random.default_rng(5)
    N_s, K, n_cols = 4, 3, 7
    ids = rng.integers(0, n_cols, size=(N_s, K))
    ids[0, 0] = ids[1, 2] = 6                    # a shared node
    blk = rng.standard_normal((5, 3 * K * N_s))
    P = np.zeros((3 * K * N_s, 3 * n_cols))
    for s in range(N_s):
        for k in range(K):
            for j in range(3):
                P[3 * (K * s + k) + j, 3 * ids[s, k] + j] = 1.0
    assert np.allclose(kb.condense_columns(blk, ids, n_cols), blk @ P, atol=1e-14)
    out = np.ones((5, 3 * n_cols))
    kb.condense_columns(blk, ids, n_cols, out=out)
    assert np.allclose(out, 1.0 + blk @ P, atol=1e-14)
    ident = kb.identity_ids(1, N_s)
    assert np.array_equal(kb.condense_columns(blk, ident, K * N_s), blk)
    with pytest.raises(ValueError):
        kb.condense_columns(blk[:, :-3], ids, n_cols)
    with pytest.raises(ValueError):
        kb.condense_columns(blk, ids, 5)             # ids reach 6


def test_mesh_arrays_memo():
    from mbem_amd.mesh import TriMesh
    calls = []

    class Counting(TriMesh):
        def centroids(self):
            calls.append(1)
            return super().centroids()

    m = Counting(vertices=np.eye(3), triangles=np.array([[0, 1, 2]]))
    memo = kb.MeshArrays()
    a = memo.field_points(m)
    b = memo.field_points(m)
    assert a is b and len(calls) == 1
    v, n = memo.source_arrays(m)
    assert v.shape == (1, 3, 3) and memo.source_arrays(m)[0] is v
    pts = np.zeros((2, 3))
    assert np.array_equal(memo.field_points(pts), pts)
    # the traction rows' kind and ids table are checked before any device call
    from mbem_amd.mesh import ElasticMaterial
    mat = ElasticMaterial(mu=1.0, lam=1.0)
    with pytest.raises(KeyError):
        kb.assemble_traction_matrix(pts, pts, m, mat, 0.1, kind="bogus")
    with pytest.raises(ValueError, match="ids shape"):
        kb.assemble_traction_matrix(pts, pts, m, mat, 0.1, order=1, ids=np.zeros((1, 1)), n_cols=1)


def test_source_geometry_matches_the_separate_calls_bit_for_bit():
    """kb._source_geometry (one gather) against _source_arrays plus the
    patch's collocation points, on every patch of a box with a dipping
    fault: equal arrays, not merely close."""
    from mbem_amd.backends.dense import _centroid_collocated
    from mbem_amd.cases import box_fault_model
    from mbem_amd.kernels import basis as kb
    model, _, _ = box_fault_model(8, fault_cells=(6, 10))
    patches = [p for r in model.regions for p in list(r.patches) + list(r.faults)]
    assert patches and all(_centroid_collocated(p) for p in patches)
    for p in patches:
        tv, nrm, x = kb._source_geometry(p.mesh)
        tv0, nrm0 = kb._source_arrays(p.mesh)
        assert np.array_equal(tv, tv0) and np.array_equal(nrm, nrm0)
        assert np.array_equal(x, np.ascontiguousarray(p.collocation_points()))
        assert tv.flags.c_contiguous and nrm.flags.c_contiguous and x.flags.c_contiguous


def test_tri_geometry_on_irregular_triangles_and_bad_indices():
    """mbem_tri_geometry (host code of the library) against numpy on random
    triangles spanning ten decades of size: equal arrays; an index outside
    the vertex list is refused before anything is written."""
    from mbem_amd import _lib
    from mbem_amd.kernels import basis as kb
    from mbem_amd.mesh import TriMesh
    rng = np.random.default_rng(7)
    verts = rng.standard_normal((300, 3)) * np.exp(rng.uniform(-12, 12, (300, 1)))
    tris = np.array([rng.choice(300, 3, replace=False) for _ in range(500)])
    mesh = TriMesh(verts, tris)
    tv, nrm, cen = kb._source_geometry(mesh)
    tv0, nrm0 = kb._source_arrays(mesh)
    assert np.array_equal(tv, tv0) and np.array_equal(nrm, nrm0)
    assert np.array_equal(cen, mesh.centroids())
    bad = np.ascontiguousarray(tris, dtype=np.int64)
    bad[3, 1] = 300
    out = np.full((500, 3), 7.0)
    st = _lib.load().mbem_tri_geometry(_lib.host_ptr(np.ascontiguousarray(verts)), 300,
                                        _lib.host_ptr(bad), 500, _lib.host_ptr(np.empty((500, 9))),
                                        _lib.host_ptr(out), _lib.host_ptr(np.empty((500, 3))))
    assert st == _lib.MBEM_EINVAL and np.all(out == 7.0)
