"""Near-face readout: a continuous displacement recovered from the
element-constant one a solve returns (mbem's src/mbem/recover.py).

A point within about half an element of a boundary patch resolves the
element edges of a piecewise-constant double layer one by one (each edge
step is a small dislocation line).  ``NearField`` holds, for every order-0
boundary patch of a model, a sparse operator from the element values to a
continuous P2 density (a value per shared vertex and edge midpoint: the
quadratic that best fits the elements two rings around the node, not across
a fault and not across a crease), and the readout reads the near pairs from
that density at a narrowed mollification (``READOUT_EPS_FACTOR``) while the
traction stays as solved.  All of this is host-side sparse bookkeeping
(numpy / scipy) over the mesh; the near pairs are then evaluated by the
device nodal kernels (``mbem_amd.evaluate``).
"""
from __future__ import annotations

import numpy as np

from . import defaults
from .model.core import FAULT_KINDS


def split_nodes(vertices, triangles, faults=(), on_fault: float = None):
    """``(nodes (M, 3), index (T, 3))``: one node per group of coincident
    vertices, split by side where it lies on one of ``faults`` (within
    ON_FAULT_OVER_EDGE of the shortest edge; the side of an element is that
    of its centroid), and the node of every element corner
    (model/shared_nodes.py:42-95)."""
    from scipy import sparse
    from scipy.spatial import cKDTree

    from .geometry import distance_to_mesh

    v = np.asarray(vertices, float)
    tri = np.asarray(triangles, np.int64)
    corners = v[tri]
    cen = corners.mean(axis=1)
    edge = np.linalg.norm(corners - np.roll(corners, 1, axis=1), axis=2)
    tol = (defaults.ON_FAULT_OVER_EDGE if on_fault is None else on_fault) * edge.min()
    pairs = cKDTree(v).query_pairs(tol, output_type="ndarray")
    if pairs.size:
        _n, label = sparse.csgraph.connected_components(
            sparse.coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])),
                              shape=(v.shape[0],) * 2), directed=False)
        tri = label[tri]
    side = np.zeros(tri.shape, np.int64)
    for f in faults:
        d, near = distance_to_mesh(v, f)
        on = d < tol
        if not on.any():
            continue
        fv = np.asarray(f.vertices, float)[np.asarray(f.triangles)]
        fn = np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0])
        for c in range(3):
            sel = on[tri[:, c]]
            k = near[tri[sel, c]]
            s = np.sign(np.einsum("ij,ij->i", cen[sel] - fv[k, 0], fn[k]))
            side[sel, c] = 3 * side[sel, c] + s.astype(np.int64) + 2
    _keys, index = np.unique(np.column_stack([tri.ravel(), side.ravel()]), axis=0,
                             return_inverse=True)
    index = np.asarray(index).ravel()
    nodes = np.empty((index.max() + 1, 3))
    nodes[index] = corners.reshape(-1, 3)
    return nodes, index.reshape(tri.shape[0], 3)


def _fit(x, hint, ptr, cols, cen, nrm, order):
    """Weights of the fitted value at each point ``x`` over its elements (CSR
    ``ptr`` / ``cols``): ``(rows, cols, vals, done)`` (recover.py:39-93).

    The fit is in the tangent plane of ``hint`` over the elements facing its
    way (a crease is not fitted across); order 2 fits a quadratic, 1 a
    plane, 0 the mean.  A point with too few elements for the order, or whose
    least-squares weights do not reproduce a constant, is left not ``done``
    for a lower order."""
    n = x.shape[0]
    row = np.repeat(np.arange(n), np.diff(ptr))
    keep = np.einsum("ij,ij->i", nrm[cols], hint[row]) > defaults.READOUT_CREASE_COS
    few = np.bincount(row, weights=keep, minlength=n) < 3
    keep |= few[row]
    row, cols = row[keep], cols[keep]
    count = np.bincount(row, minlength=n)
    need = (1, defaults.READOUT_MIN_LINEAR, defaults.READOUT_MIN_QUADRATIC)[order]
    done = count >= need
    # an orthonormal tangent frame (a, b) per point
    a = np.cross(hint, [1.0, 0.0, 0.0])
    alt = np.linalg.norm(a, axis=1) < 0.3
    a[alt] = np.cross(hint[alt], [0.0, 1.0, 0.0])
    a /= np.linalg.norm(a, axis=1)[:, None]
    b = np.cross(hint, a)
    d = cen[cols] - x[row]
    p = np.einsum("ij,ij->i", d, a[row])
    q = np.einsum("ij,ij->i", d, b[row])
    start = np.concatenate([[0], np.cumsum(count)])
    out_r, out_c, out_v = [], [], []
    for k in np.unique(count[done]):
        sel = np.flatnonzero(done & (count == k))
        at = start[sel][:, None] + np.arange(k)[None, :]
        pk, qk = p[at], q[at]
        if order == 0:
            w = np.full(at.shape, 1.0 / k)
        else:
            s = np.maximum(np.abs(pk).max(axis=1), np.abs(qk).max(axis=1))
            s = np.where(s > 0.0, s, 1.0)[:, None]
            pk, qk = pk / s, qk / s
            basis = [np.ones_like(pk), pk, qk] + ([pk * pk, pk * qk, qk * qk] if order == 2 else [])
            # the fitted value at the point (p = q = 0) is the constant term
            w = np.linalg.pinv(np.stack(basis, axis=2), rcond=1e-8)[:, 0, :]
            bad = np.abs(w.sum(axis=1) - 1.0) > 1e-8
            done[sel[bad]] = False
            sel, at, w = sel[~bad], at[~bad], w[~bad]
        out_r.append(np.repeat(sel, k))
        out_c.append(cols[at].ravel())
        out_v.append(w.ravel())

    def cat(parts, dt):
        return np.concatenate(parts) if parts else np.zeros(0, dt)
    return cat(out_r, np.int64), cat(out_c, np.int64), cat(out_v, float), done


def recovery_operator(mesh, faults=()):
    """Sparse (6 N, N): element values -> the element-major nodal values of
    a continuous P2 density on ``mesh`` (node order v1 v2 v3 m12 m23 m31),
    one value per shared node, split across ``faults`` (recover.py:96-168).
    Rows sum to one; a quadratic field is returned exactly wherever two rings
    of elements determine one, and lower orders / wider rings are used only
    where they do not."""
    from scipy import sparse

    V = np.asarray(mesh.vertices, float)
    T = np.asarray(mesh.triangles, np.int64)
    N = T.shape[0]
    nodes, idx = split_nodes(V, T, faults)
    M = nodes.shape[0]
    cen = V[T].mean(axis=1)
    nrm, _ = mesh.normals_and_areas()
    nrm = np.asarray(nrm, float)
    # one node per edge of the split surface, in each element's edge order
    ea, eb = idx, idx[:, [1, 2, 0]]
    key = np.minimum(ea, eb) * np.int64(M) + np.maximum(ea, eb)
    _u, first, edge = np.unique(key.ravel(), return_index=True, return_inverse=True)
    edge = np.asarray(edge).reshape(N, 3)
    E = first.size
    fe, fk = np.divmod(first, 3)
    mid = 0.5 * (V[T[fe, fk]] + V[T[fe, (fk + 1) % 3]])
    x = np.vstack([nodes, mid])
    # a normal per fit point: that of one element owning it
    owner = np.empty(M + E, np.int64)
    owner[idx.ravel()[::-1]] = np.repeat(np.arange(N), 3)[::-1]
    owner[M:] = fe
    hint = nrm[owner]
    e2n = sparse.csr_matrix((np.ones(3 * N), (np.repeat(np.arange(N), 3), idx.ravel())),
                            shape=(N, M))
    n2e = e2n.T.tocsr()
    e2e = (e2n @ n2e).tocsr()
    pick = sparse.csr_matrix((np.ones(2 * E), (np.repeat(np.arange(E), 2),
                                               np.column_stack([ea.ravel()[first],
                                                                eb.ravel()[first]]).ravel())),
                             shape=(E, M))
    one = sparse.vstack([n2e, pick @ n2e]).tocsr()
    rows, cols, vals = [], [], []
    todo = np.arange(M + E)
    for order, rings in ((2, 2), (2, 3), (2, 4), (1, 2), (1, 3), (0, 1)):
        if todo.size == 0:
            break
        reach = one[todo]
        for _ in range(rings - 1):
            reach = reach @ e2e
        reach = reach.tocsr()
        reach.sort_indices()
        r, c, v, done = _fit(x[todo], hint[todo], reach.indptr, reach.indices, cen, nrm, order)
        rows.append(todo[r])
        cols.append(c)
        vals.append(v)
        todo = todo[~done]
    if todo.size:
        raise ValueError(f"{todo.size} node(s) of the surface belong to no element")
    W = sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(M + E, N))
    ids = np.concatenate([idx, M + edge], axis=1)
    return W[ids.ravel()].tocsr()


class NearField:
    """What ``evaluate_displacement`` / ``evaluate_stress`` take as ``near=``:
    the recovery operator of every order-0 boundary patch of ``model``, built
    once for its geometry (recover.py:171-203).  Patches are identified by
    identity, so a state's model that shares its base's patches shares this."""

    def __init__(self, model, eps_factor: float = None, radius_h: float = None,
                 standoff_h: float = None):
        pick = lambda v, d: d if v is None else float(v)  # noqa: E731
        self.eps_factor = pick(eps_factor, defaults.READOUT_EPS_FACTOR)
        self.radius_h = pick(radius_h, defaults.READOUT_NEAR_RADIUS_H)
        self.standoff_h = pick(standoff_h, defaults.READOUT_STANDOFF_H)
        faults = [f.mesh for r in model.regions for f in r.faults]
        self._ops = {}
        self._patches = []
        for r in model.regions:
            for p in r.patches:
                if id(p) in self._ops or getattr(p, "order", 0) != 0 or p.bc in FAULT_KINDS:
                    continue
                self._ops[id(p)] = recovery_operator(p.mesh, faults)
                self._patches.append(p)

    def recovers(self, patch) -> bool:
        return id(patch) in self._ops

    def recover(self, patch, values) -> np.ndarray:
        """Element values (N, 3) of ``patch`` -> the P2 nodal density (6 N, 3)."""
        return np.ascontiguousarray(self._ops[id(patch)] @ np.asarray(values, float))
