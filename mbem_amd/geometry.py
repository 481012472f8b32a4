"""Exact-geometry helpers for the model layer and the readout warnings
(mbem's src/mbem/geometry.py): signed solid angles of a mesh at points, and
point-to-mesh distances.  Host-side numpy; not on the kernel path.
"""
from __future__ import annotations

import numpy as np


def solid_angle_batch(mesh, points, chunk: int = 64) -> np.ndarray:
    """(N,) signed solid angle of ``mesh`` at each point: the van Oosterom
    formula 2 atan2(a.(b x c), |a||b||c| + |c| a.b + |a| b.c + |b| a.c) summed
    over triangles (geometry.py:8-56), ``chunk`` points at a time, written
    out in components over (point, triangle) arrays."""
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    tv = np.asarray(mesh.vertices, float)[np.asarray(mesh.triangles)]
    va, vb, vc = (np.ascontiguousarray(tv[:, k].T) for k in range(3))      # (3, N_tri) each
    out = np.empty(pts.shape[0])
    for p0 in range(0, pts.shape[0], chunk):
        p = pts[p0:p0 + chunk]
        ax, ay, az = (va[k][None, :] - p[:, k:k + 1] for k in range(3))
        bx, by, bz = (vb[k][None, :] - p[:, k:k + 1] for k in range(3))
        cx, cy, cz = (vc[k][None, :] - p[:, k:k + 1] for k in range(3))
        ra = np.sqrt(ax * ax + ay * ay + az * az)
        rb = np.sqrt(bx * bx + by * by + bz * bz)
        rc = np.sqrt(cx * cx + cy * cy + cz * cz)
        numer = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx)
        denom = (ra * rb * rc + rc * (ax * bx + ay * by + az * bz)
                 + ra * (bx * cx + by * cy + bz * cz) + rb * (ax * cx + ay * cy + az * cz))
        out[p0:p0 + chunk] = (2.0 * np.arctan2(numer, denom)).sum(axis=1)
    return out


def point_triangle_distance(p, a, b, c):
    """Exact distance from points to triangles (broadcastable (..., 3)): the
    closest point of a triangle by its Voronoi regions (vertex, edge,
    interior) -- the standard closest-point construction, as
    geometry.py:59-87 uses."""
    ab, ac = b - a, c - a
    ap, bp, cp = p - a, p - b, p - c
    d1, d2 = (ab * ap).sum(-1), (ac * ap).sum(-1)
    d3, d4 = (ab * bp).sum(-1), (ac * bp).sum(-1)
    d5, d6 = (ab * cp).sum(-1), (ac * cp).sum(-1)
    va, vb, vc = d3 * d6 - d5 * d4, d5 * d2 - d1 * d6, d1 * d4 - d3 * d2
    with np.errstate(divide="ignore", invalid="ignore"):
        s = va + vb + vc
        q = a + ab * (vb / s)[..., None] + ac * (vc / s)[..., None]
        t = (d1 / (d1 - d3))[..., None]
        q = np.where(((vc <= 0) & (d1 >= 0) & (d3 <= 0))[..., None], a + ab * t, q)
        t = (d2 / (d2 - d6))[..., None]
        q = np.where(((vb <= 0) & (d2 >= 0) & (d6 <= 0))[..., None], a + ac * t, q)
        t = ((d4 - d3) / ((d4 - d3) + (d5 - d6)))[..., None]
        q = np.where(((va <= 0) & (d4 >= d3) & (d5 >= d6))[..., None], b + (c - b) * t, q)
    q = np.where(((d1 <= 0) & (d2 <= 0))[..., None], a, q)
    q = np.where(((d3 >= 0) & (d4 <= d3))[..., None], b, q)
    q = np.where(((d6 >= 0) & (d5 <= d6))[..., None], c, q)
    return np.linalg.norm(p - q, axis=-1)


def distance_to_mesh(points, mesh, k: int = 32):
    """(d, idx): exact distance from each point to ``mesh`` over the ``k``
    triangles with the nearest centroids (geometry.py:90-105)."""
    from scipy.spatial import cKDTree

    points = np.asarray(points, float).reshape(-1, 3)
    tv = np.asarray(mesh.vertices, float)[np.asarray(mesh.triangles)]
    k = min(k, tv.shape[0])
    _, idx = cKDTree(mesh.centroids()).query(points, k=k)
    idx = np.asarray(idx).reshape(points.shape[0], k)
    d = point_triangle_distance(points[:, None, :], tv[idx, 0], tv[idx, 1], tv[idx, 2])
    j = d.argmin(axis=1)
    rows = np.arange(points.shape[0])
    return d[rows, j], idx[rows, j]
