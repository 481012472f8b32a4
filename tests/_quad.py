"""Independent references for the closed-form pair integrals: Gauss product
quadrature of the POINTWISE mollified Kelvin kernels, with the derivatives
taken symbolically (sympy) from the displacement kernel itself

    G_ij(d) = C1 [ (3-4nu) d_ij / R + d_i d_j / R^3 + 2(1-nu) eps^2 d_ij / R^3 ],
    R = sqrt(|d|^2 + eps^2),  C1 = 1/(16 pi mu (1-nu)),  d = x - y

(mbem's src/clq/pointwise.py:1-20 states the same point kernel), so nothing
here shares arithmetic with the closed forms under test: the T, Kelvin-stress
and dislocation-stress kernels below are built from symbolic first and second
derivatives of G, contracted with the traction operator
``T_ij = -[lam n_j dG_im/dx_m + mu n_m dG_ij/dx_m + mu n_k dG_ik/dx_j]``
(tri_kernels.py:24-26) and Hooke's law.  The mollified integrands are smooth
(R >= eps), so a collapsed Gauss rule on a uniformly subdivided triangle
converges geometrically; every call returns the result at two subdivision
depths so a test can assert convergence before using it.
"""
from __future__ import annotations

import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def _symbolic():
    import sympy as sp

    d = sp.symbols("d0:3", real=True)
    eps, nu = sp.symbols("eps nu", positive=True)
    R = sp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + eps ** 2)
    c34 = 3 - 4 * nu
    G = [[(c34 * sp.KroneckerDelta(i, j) / R + d[i] * d[j] / R ** 3
           + 2 * (1 - nu) * eps ** 2 * sp.KroneckerDelta(i, j) / R ** 3)
          for j in range(3)] for i in range(3)]
    # without the C1 prefactor; d/dx = d/dd because d = x - y
    dG = [[[sp.diff(G[i][j], d[m]) for m in range(3)] for j in range(3)] for i in range(3)]
    d2G = [[[[sp.diff(dG[i][j][m], d[q]) for q in range(3)] for m in range(3)]
             for j in range(3)] for i in range(3)]
    args = (d[0], d[1], d[2], eps, nu)
    fG = sp.lambdify(args, [G[i][j] for i in range(3) for j in range(3)], "numpy")
    fdG = sp.lambdify(args, [dG[i][j][m] for i in range(3) for j in range(3) for m in range(3)],
                      "numpy")
    fd2G = sp.lambdify(args, [d2G[i][j][m][q] for i in range(3) for j in range(3)
                              for m in range(3) for q in range(3)], "numpy")
    return fG, fdG, fd2G


def _eval(f, shape, d, eps, nu):
    vals = f(d[:, 0], d[:, 1], d[:, 2], eps, nu)
    n = d.shape[0]
    out = np.empty((n,) + shape)
    flat = out.reshape(n, -1)
    for k, v in enumerate(vals):
        flat[:, k] = v
    return out


def _gauss_triangle(n):
    """Collapsed (Duffy) Gauss product rule on the unit triangle:
    barycentric points (n*n, 2) in (s, t) with weights summing to 1/2."""
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    u, v = np.meshgrid(x, x, indexing="ij")
    wu, wv = np.meshgrid(w, w, indexing="ij")
    s = u.ravel()
    t = (v * (1.0 - u)).ravel()
    wt = (wu * wv * (1.0 - u)).ravel()
    return s, t, wt


def _subdivide(tv, depth):
    tris = [np.asarray(tv, float)]
    for _ in range(depth):
        nxt = []
        for a, b, c in tris:
            ab, bc, ca = 0.5 * (a + b), 0.5 * (b + c), 0.5 * (c + a)
            nxt += [np.array([a, ab, ca]), np.array([ab, b, bc]),
                    np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nxt
    return np.array(tris)


def _points(tv, depth, n):
    s, t, wt = _gauss_triangle(n)
    sub = _subdivide(tv, depth)
    a, b, c = sub[:, 0], sub[:, 1], sub[:, 2]
    area2 = np.linalg.norm(np.cross(b - a, c - a), axis=1)
    y = (a[:, None, :] + s[None, :, None] * (b - a)[:, None, :]
         + t[None, :, None] * (c - a)[:, None, :]).reshape(-1, 3)
    w = (area2[:, None] * wt[None, :]).ravel()
    return y, w


def _wsum(w, vals):
    """sum_q w_q vals[q] with exactly rounded (math.fsum) accumulation, so the
    quadrature's own rounding stays far below the closed forms' accuracy."""
    import math
    prod = (w.reshape((-1,) + (1,) * (vals.ndim - 1)) * vals).reshape(vals.shape[0], -1)
    out = np.array([math.fsum(prod[:, k]) for k in range(prod.shape[1])])
    return out.reshape(vals.shape[1:])


def lagrange_weight(tv, order, k):
    """y (q, 3) -> N_k(y): the textbook Lagrange basis of the triangle tv in
    barycentric coordinates (P1: lam_k; P2: vertices lam_i (2 lam_i - 1),
    edge midpoints 4 lam_i lam_j on edges 12, 23, 31), written independently
    of the node-lattice product formula of the kernels under test."""
    tv = np.asarray(tv, float)
    T = np.stack([tv[1] - tv[0], tv[2] - tv[0]], axis=1)      # (3, 2)

    def N(y):
        st = np.linalg.lstsq(T, (y - tv[0]).T, rcond=None)[0].T   # (q, 2)
        lam = np.column_stack([1.0 - st[:, 0] - st[:, 1], st[:, 0], st[:, 1]])
        if order == 0:
            return np.ones(y.shape[0])
        if order == 1:
            return lam[:, k]
        if k < 3:
            return lam[:, k] * (2.0 * lam[:, k] - 1.0)
        i, j = ((0, 1), (1, 2), (2, 0))[k - 3]
        return 4.0 * lam[:, i] * lam[:, j]
    return N


def _integrate(kind, tv, obs, eps, mu, lam, normal, depth, n, weight=None):
    nu = lam / (2.0 * (lam + mu))
    C1 = 1.0 / (16.0 * np.pi * mu * (1.0 - nu))
    fG, fdG, fd2G = _symbolic()
    y, w = _points(tv, depth, n)
    if weight is not None:
        w = w * weight(y)
    d = np.asarray(obs, float)[None, :] - y
    nrm = np.asarray(normal, float)
    if kind == "G":
        return C1 * _wsum(w, _eval(fG, (3, 3), d, eps, nu))
    if kind in ("T", "S"):
        P = C1 * _wsum(w, _eval(fdG, (3, 3, 3), d, eps, nu))
        if kind == "T":
            # T_ij = -[lam n_j dG_im/dx_m + mu n_m dG_ij/dx_m + mu n_k dG_ik/dx_j]
            tr = np.einsum("imm->i", P)
            return -(lam * np.einsum("i,j->ij", tr, nrm) + mu * np.einsum("ijm,m->ij", P, nrm)
                     + mu * np.einsum("ikj,k->ij", P, nrm))
        # force -> stress S[i,j,k]: sigma_ij = lam d_ij eps_mm + 2 mu eps_ij, u_i = G_ik f_k
        trk = np.einsum("mkm->k", P)
        return (lam * np.einsum("ij,k->ijk", np.eye(3), trk)
                + mu * (np.einsum("ikj->ijk", P) + np.einsum("jki->ijk", P)))
    if kind == "H":
        # ID2G[r,p,s,q] = d2 G_rp / dx_s dx_q (lambdify order: G[r][p], d/ds, d/dq)
        ID2G = C1 * _wsum(w, _eval(fd2G, (3, 3, 3, 3), d, eps, nu))
        tr = np.einsum("rpsp->rs", ID2G)
        B = (lam * np.einsum("k,rs->rsk", nrm, tr) + mu * np.einsum("q,rksq->rsk", nrm, ID2G)
             + mu * np.einsum("p,rpsk->rsk", nrm, ID2G))
        trB = np.einsum("rrk->k", B)
        return -(lam * np.einsum("mn,k->mnk", np.eye(3), trB)
                 + mu * (B + np.einsum("nmk->mnk", B)))
    if kind == "E":
        R2 = np.einsum("qi,qi->q", d, d) + eps ** 2
        phi = float(_wsum(w, 15.0 * eps ** 4 / (8.0 * np.pi * R2 ** 3.5)))
        e = np.eye(3)
        return phi * (lam * np.einsum("mn,k->mnk", e, nrm)
                      + mu * (np.einsum("mk,n->mnk", e, nrm) + np.einsum("nk,m->mnk", e, nrm)))
    if kind == "I":
        # the scalar moments I_n = int 1/R^n for n = 1, 3, 5, 7
        R2 = np.einsum("qi,qi->q", d, d) + eps ** 2
        return np.array([float(_wsum(w, 1.0 / R2 ** (k / 2.0))) for k in (1, 3, 5, 7)])
    raise ValueError(kind)


def pair(kind, tv, obs, eps, mu=1.0, lam=1.0, normal=None, depth=3, n=16, weight=None):
    """Quadrature of one pair at ``depth`` and ``depth + 1`` subdivisions.

    kind: "G" (3,3) force->displacement; "T" (3,3) slip->displacement with
    the source normal; "S" (3,3,3) force->stress; "H" (3,3,3) slip->total
    stress; "E" (3,3,3) eigenstress; "I" (4,) scalar moments I1..I7.
    ``weight(y)``, if given, multiplies the integrand (a nodal shape function).
    """
    tv = np.asarray(tv, float)
    if normal is None:
        c = np.cross(tv[1] - tv[0], tv[2] - tv[0])
        normal = c / np.linalg.norm(c)
    a = _integrate(kind, tv, obs, eps, mu, lam, normal, depth, n, weight)
    b = _integrate(kind, tv, obs, eps, mu, lam, normal, depth + 1, n, weight)
    return a, b
