"""Cluster tree over element centroids (mbem's src/mbem/la/cluster.py:19-78):
binary principal-axis bisection at the median projection, leaves of at most
``min_leaf`` elements.  Host-side numpy; it decides the block-Jacobi chunks.
"""
from __future__ import annotations

import numpy as np

from .. import defaults


class ClusterNode:
    """A node: its element indices and children; the bounding box (lo, hi)
    is computed when read (the block-Jacobi chunks need only the leaves)."""
    __slots__ = ("indices", "_centroids", "left", "right")

    def __init__(self, indices, centroids):
        self.indices = indices
        self._centroids = centroids
        self.left = None
        self.right = None

    @property
    def lo(self):
        return self._centroids[self.indices].min(axis=0)

    @property
    def hi(self):
        return self._centroids[self.indices].max(axis=0)

    @property
    def is_leaf(self) -> bool:
        return self.left is None

    @property
    def n(self) -> int:
        return len(self.indices)


def build_cluster_tree(centroids: np.ndarray, min_leaf: int = defaults.CLUSTER_MIN_LEAF):
    """cluster.py:56-76: split along the principal axis of the covariance at the
    median projection (``proj < median`` goes left); a split that leaves one
    side empty stops."""
    centroids = np.asarray(centroids, float)

    def _build(idx):
        node = ClusterNode(idx, centroids)
        if len(idx) <= min_leaf:
            return node
        c = centroids[idx]
        c = c - c.mean(axis=0)
        _, vecs = np.linalg.eigh(c.T @ c)
        proj = c @ vecs[:, -1]
        left = proj < np.median(proj)
        nl = int(np.count_nonzero(left))
        if nl == 0 or nl == len(idx):
            return node
        node.left = _build(idx[left])
        node.right = _build(idx[~left])
        return node

    return _build(np.arange(centroids.shape[0]))


def leaves(node) -> list:
    """Leaf index sets, left to right."""
    out, stack = [], [node]
    while stack:
        nd = stack.pop()
        if nd.is_leaf:
            out.append(nd.indices)
        else:
            stack.append(nd.right)
            stack.append(nd.left)
    return out
