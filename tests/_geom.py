"""Seeded test geometries: random and adversarial source triangles with
observation points on, near and far from them (tests only)."""
from __future__ import annotations

import numpy as np


def random_triangles(rng, n, scale=1.0, needle=False):
    tv = rng.normal(size=(n, 3, 3)) * scale
    if needle:
        # long thin elements, height / L ~ 1e-3
        base = rng.normal(size=(n, 1, 3)) * scale
        d = rng.normal(size=(n, 1, 3))
        d /= np.linalg.norm(d, axis=2, keepdims=True)
        p = rng.normal(size=(n, 1, 3))
        p -= np.sum(p * d, axis=2, keepdims=True) * d
        p /= np.linalg.norm(p, axis=2, keepdims=True)
        tv = np.concatenate([base, base + 2.0 * scale * d, base + scale * d + 2e-3 * scale * p], axis=1)
    return np.ascontiguousarray(tv)


def normals_of(tv):
    c = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    return c / np.linalg.norm(c, axis=1, keepdims=True)


def adversarial_points(rng, tv, n_far=8):
    """Centroids, lifted centroids, edge midpoints, vertices, in-plane
    exterior points and far points for every triangle in ``tv``."""
    nrm = normals_of(tv)
    pts = []
    for t, n in zip(tv, nrm):
        c = t.mean(axis=0)
        L = np.linalg.norm(t[1] - t[0])
        pts += [c, c + 1e-3 * L * n, c - 0.3 * L * n, 0.5 * (t[0] + t[1]),
                0.5 * (t[1] + t[2]) + 1e-6 * L * n, t[2], t[0] + 1e-9 * L * n,
                t[0] + 2.0 * (t[1] - t[0]),           # in-plane, on the edge line outside
                c + 3.0 * (c - t[2])]                  # in-plane exterior
    pts += list(rng.normal(size=(n_far, 3)) * 30.0)
    return np.ascontiguousarray(np.array(pts))


def materials():
    # (mu, lam): nu = 0.25, 0.30 and a near-incompressible 0.45
    return [(30.0, 30.0), (30.0, 45.0), (1.0, 9.0)]
