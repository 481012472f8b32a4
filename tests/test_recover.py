"""The near-face recovery operator (mbem_amd/recover.py, after recover.py:39-203)
pinned on the CPU by what it must do to fields whose answer is known:

* rows sum to one (a constant comes back exactly);
* a linear field sampled at the element centroids comes back exactly at
  every node of a flat patch, and a quadratic one at every node whose fit
  could be quadratic (all nodes away from the patch's rim);
* one value per shared node: the rows of coincident nodes of neighbouring
  elements are identical, so the recovered density is continuous;
* across a fault the nodes are split and nothing is fitted across: a field
  linear on each side with a jump at the fault comes back exactly on both
  sides, the on-fault nodes with their own side's value;
* a crease is not fitted across: on the box's sides (four panels at right
  angles) a field linear on each panel separately comes back exactly;
* NearField recovers the order-0 boundary patches of a model and not its
  faults.
"""
import numpy as np
import pytest


def _nodes(mesh):
    """Element-major node coordinates (6 N, 3): v1 v2 v3 m12 m23 m31."""
    v = np.asarray(mesh.vertices, float)[np.asarray(mesh.triangles)]
    mids = 0.5 * (v + v[:, [1, 2, 0]])
    return np.concatenate([v, mids], axis=1).reshape(-1, 3)


def _plate(n=12, half=20.0):
    from mbem_amd.cases import rectangular_patch
    return rectangular_patch((-half, half), (-half, half), 0.0, n, n, normal_up=True)


def test_constant_linear_quadratic_on_a_flat_patch():
    from mbem_amd.recover import recovery_operator
    mesh = _plate()
    W = recovery_operator(mesh)
    N = mesh.n_triangles
    assert W.shape == (6 * N, N)
    assert np.allclose(W @ np.ones(N), 1.0, rtol=0, atol=1e-12)
    cen, nodes = mesh.centroids(), _nodes(mesh)
    lin = lambda p: 0.3 + 0.02 * p[:, 0] - 0.05 * p[:, 1]  # noqa: E731
    assert np.abs(W @ lin(cen) - lin(nodes)).max() < 1e-12
    quad = lambda p: 0.3 + 0.02 * p[:, 0] - 0.05 * p[:, 1] + 1e-3 * (p[:, 0] ** 2 - 2 * p[:, 0] * p[:, 1] + 0.5 * p[:, 1] ** 2)  # noqa: E731,E501
    err = np.abs(W @ quad(cen) - quad(nodes))
    inner = np.all(np.abs(nodes) < 20.0 - 2 * 40.0 / 12 - 1e-9, axis=1)
    assert inner.sum() > 0.3 * inner.size
    assert err[inner].max() < 1e-10
    # near the rim a lower-order fit may take over, but never worse than a plane
    assert err.max() < 0.2 * np.abs(quad(nodes) - lin(nodes)).max() + 1e-12


def test_one_value_per_shared_node():
    from mbem_amd.recover import recovery_operator, split_nodes
    mesh = _plate(8)
    W = recovery_operator(mesh).toarray()
    nodes, idx = split_nodes(mesh.vertices, mesh.triangles)
    rows = {}
    for s in range(mesh.n_triangles):
        for k in range(3):
            rows.setdefault(idx[s, k], []).append(W[6 * s + k])
    shared = [r for r in rows.values() if len(r) > 1]
    assert len(shared) > 10
    for r in shared:
        for other in r[1:]:
            assert np.array_equal(other, r[0])


def test_fault_splits_the_nodes():
    from mbem_amd.cases import vertical_panel
    from mbem_amd.recover import recovery_operator, split_nodes
    mesh = _plate(12)
    fault = vertical_panel("x", 0.0, (-20.0, 20.0), (-5.0, 5.0), 12, 3, +1)
    nodes0, _ = split_nodes(mesh.vertices, mesh.triangles)
    nodes1, idx1 = split_nodes(mesh.vertices, mesh.triangles, [fault])
    on_line = np.sum(np.abs(np.asarray(mesh.vertices)[:, 0]) < 1e-9)
    assert nodes1.shape[0] == nodes0.shape[0] + on_line
    W = recovery_operator(mesh, [fault])
    cen, nodes = mesh.centroids(), _nodes(mesh)
    side_e = np.sign(cen[:, 0])
    f = lambda p, s: np.where(s > 0, 1.0 + 0.01 * p[:, 1], -2.0 + 0.03 * p[:, 0])  # noqa: E731
    got = W @ f(cen, side_e)
    want = f(nodes, np.repeat(side_e, 6))
    assert np.abs(got - want).max() < 1e-11


def test_crease_is_not_fitted_across():
    from mbem_amd.cases import box_meshes
    from mbem_amd.recover import recovery_operator
    _, sides, _ = box_meshes(8)
    W = recovery_operator(sides)
    cen, nodes = sides.centroids(), _nodes(sides)
    nrm, _ = sides.normals_and_areas()
    # a field linear on each panel, different panel to panel
    g = {(1, 0): np.array([0.0, 0.01, 0.02]), (-1, 0): np.array([0.0, -0.03, 0.01]),
         (0, 1): np.array([0.02, 0.0, -0.01]), (0, -1): np.array([0.01, 0.0, 0.04])}

    def field(p, n):
        out = np.empty(p.shape[0])
        for (a, b), grad in g.items():
            sel = (np.round(n[:, 0]) == a) & (np.round(n[:, 1]) == b)
            out[sel] = 1.0 + (a + 2 * b) + p[sel] @ grad
        return out
    got = W @ field(cen, nrm)
    want = field(nodes, np.repeat(nrm, 6, axis=0))
    # the box's vertical edges are shared by two panels at right angles: a node
    # there belongs to one fit, of its owner's panel, which need not be the
    # element's -- compare away from those edges
    edge = (np.abs(np.abs(nodes[:, 0]) - 20.0) < 1e-9) & (np.abs(np.abs(nodes[:, 1]) - 20.0) < 1e-9)
    assert np.abs(got - want)[~edge].max() < 1e-10


def test_near_field_covers_boundary_patches_only():
    from mbem_amd.cases import box_fault_model
    from mbem_amd.recover import NearField
    model, _, _ = box_fault_model(6, fault_cells=(4, 6))
    near = NearField(model)
    region = model.regions[0]
    assert all(near.recovers(p) for p in region.patches)
    assert not any(near.recovers(f) for f in region.faults)
    p = region.patches[0]
    u = np.random.default_rng(0).normal(size=(p.mesh.n_triangles, 3))
    assert near.recover(p, u).shape == (6 * p.mesh.n_triangles, 3)
    with pytest.raises(KeyError):
        near.recover(region.faults[0], u)
