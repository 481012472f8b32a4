"""ORACLE -- test infrastructure only (never the product path).

CPU restatement of the two-level block-Jacobi preconditioner: block-Jacobi
over given chunks (explicit LU per chunk, la/preconditioner.py:421-470 with
"region" grouping) plus the coarse correction over W, the per-(chunk, slot
kind) shape functions (constants, rotations p x e_b, linears p_e e_b of the
group's centred, RMS-scaled collocation points):

    y = argmin |r - A W y|,   z = W y + M (r - A W y)

(least squares with numpy; the device path orthogonalises A W instead, which
gives the same y when A W has full column rank).

Pinned by identities (tests/test_coarse_oracle.py): z = w exactly for
r = A w with w in span W; the constants are the group indicators times e_b;
each rotation field is p x e_b; and FGMRES with the two-level preconditioner
reaches the LU solution.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def group_shapes(p, comps, basis):
    """Shape values (m, q) at DOFs with centred, scaled positions p and
    components comps, written out one DOF at a time."""
    q = {"const": 3, "rigid": 6, "linear": 12}[basis]
    out = np.zeros((p.shape[0], q))
    for t in range(p.shape[0]):
        a = int(comps[t])
        for b in range(3):
            out[t, b] = 1.0 if a == b else 0.0
        if basis == "rigid":
            for b in range(3):
                e = np.zeros(3)
                e[b] = 1.0
                out[t, 3 + b] = np.cross(p[t], e)[a]
        elif basis == "linear":
            for e in range(3):
                for b in range(3):
                    out[t, 3 + 3 * e + b] = p[t, e] if a == b else 0.0
    return out


def coarse_basis(chunk_dofs, pts, kind, comp, basis):
    """Dense W (n, K) with the columns of every (chunk, kind) group."""
    n = pts.shape[0]
    cols = []
    for dofs in chunk_dofs:
        dofs = np.asarray(dofs)
        for k in (0, 1):
            d = dofs[kind[dofs] == k]
            if d.size == 0:
                continue
            p = pts[d] - pts[d].mean(axis=0)
            s = np.sqrt(np.mean(np.sum(p * p, axis=1)))
            if s > 0:
                p = p / s
            vals = group_shapes(p, comp[d], basis)
            for j in range(vals.shape[1]):
                w = np.zeros(n)
                w[d] = vals[:, j]
                cols.append(w)
    return np.array(cols).T


def block_jacobi(A, chunk_dofs):
    """z = M r with every chunk solved exactly (its LU)."""
    facs = [(np.asarray(d), sla.lu_factor(A[np.ix_(d, d)])) for d in chunk_dofs]

    def M(r):
        z = np.zeros_like(r)
        for d, lu in facs:
            z[d] = sla.lu_solve(lu, r[d])
        return z
    return M


def two_level(A, W, M):
    C = A @ W

    def P(r):
        y, *_ = np.linalg.lstsq(C, r, rcond=None)
        return W @ y + M(r - C @ y)
    return P
