"""Cluster tree over element centroids (mbem's src/mbem/la/cluster.py:19-78):
binary principal-axis bisection at the median projection, leaves of at most
``min_leaf`` elements.  Host-side numpy; it decides the block-Jacobi chunks.
"""
from __future__ import annotations

import numpy as np

from .. import defaults


class ClusterNode:
    __slots__ = ("indices", "lo", "hi", "left", "right")

    def __init__(self, indices, centroids):
        self.indices = indices
        pts = centroids[indices]
        self.lo = pts.min(axis=0)
        self.hi = pts.max(axis=0)
        self.left = None
        self.right = None

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
        pts = centroids[idx]
        mean = pts.mean(axis=0)
        cov = (pts - mean).T @ (pts - mean)
        _, vecs = np.linalg.eigh(cov)
        proj = (pts - mean) @ vecs[:, -1]
        median = np.median(proj)
        left = proj < median
        if left.sum() == 0 or (~left).sum() == 0:
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
