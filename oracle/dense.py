"""ORACLE -- test infrastructure only (never the product path).

CPU restatement of the reference's dense system build and solves:

* ``assemble(system, eps, jump)`` -> (A, b, calib): backends/dense.py:480-528
  with the block cache (:466-478), the calibrated diagonal's row sums over
  every (q, p) T block (:495-500, model/equations.py:512-537), the block
  placement ``A[r0:r1, c0:c1] += scale * blk`` plus the free term
  (equations.py:540-582), the RHS terms ``b += scale * blk @ vector``
  (:514-523) and add_jump_rhs (equations.py:585-603).  Blocks come from the
  oracle pair kernels (oracle/tri_kernels_ref.c).
* ``lu_solve(A, b)``: backends/dense.py:532-571 with la/lu.py "double"
  (scipy LAPACK getrf/getrs).
* ``fgmres``: la/solver.py:115-245 line for line (modified Gram-Schmidt).
* ``BlockJacobiGS``: la/preconditioner.py:314-543 (patch grouping, forced
  block-Jacobi or dense-LU rung, scipy lu_factor per chunk, lower couplings
  as dense block products).

The system description (BlockSystem: terms, slots, signs) is taken as
input; every number is computed here.

Pinning: the pair kernels by tests/test_oracle_pinning.py; the assembly and
solves by manufactured solutions (rigid translation -> zero traction to
rounding, the fault-sign physics statement) and by LU == FGMRES on the same
matrix, in tests/test_oracle_system.py.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import tri_kernels as ok

KERNEL_T, KERNEL_U = "H", "G"


def _u_coeffs(mu, lam):
    # basis.py:46-55
    nu = lam / (2.0 * (lam + mu))
    one_minus_nu = (lam + 2.0 * mu) / (2.0 * (lam + mu))
    C1 = 1.0 / (16.0 * np.pi * mu * one_minus_nu)
    return np.array([(3.0 - 4.0 * nu) * C1, 1.0 / (8.0 * np.pi * mu), C1])


def _t_coeffs(mu, lam):
    # basis.py:58-72
    nu = lam / (2.0 * (lam + mu))
    one_minus_nu = (lam + 2.0 * mu) / (2.0 * (lam + mu))
    C1 = 1.0 / (16.0 * np.pi * mu * one_minus_nu)
    c34 = 3.0 - 4.0 * nu
    cL, cM = lam * C1, mu * C1
    blob = 6.0 * one_minus_nu
    return np.array([c34 * cL, -cL, blob * cL, c34 * cM, -cM, blob * cM])


def _eps_of(eps, patch):
    """basis.py:187-254 (scalar, array, dict by patch name, or "auto")."""
    from mbem_amd.kernels.basis import resolve_patch_eps
    return resolve_patch_eps(eps, patch)


def _arrays(patch):
    tv = np.ascontiguousarray(np.asarray(patch.mesh.vertices, float)[np.asarray(patch.mesh.triangles)])
    c = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    nrm = c / np.linalg.norm(c, axis=1)[:, None]
    return tv, np.ascontiguousarray(nrm)


def assemble(system, eps, jump="calibrated", materials=None):
    """(A, b, calib) of a BlockSystem on the CPU (backends/dense.py:480-528)."""
    model = system.model
    layout = system.layout
    mats = materials or {r.name: r.material for r in model.regions}
    cache = {}

    def block(q, p, kernel, region):
        m = mats[region.name]
        key = (id(q), id(p), kernel, id(m))
        if key not in cache:
            x = np.ascontiguousarray(q.collocation_points())
            tv, nrm = _arrays(p)
            e = _eps_of(eps, p)
            if kernel == KERNEL_T:
                cache[key] = ok.t_matrix_direct(x, tv, nrm, e, *_t_coeffs(m.mu, m.lam))
            else:
                cache[key] = ok.u_matrix_direct(x, tv, e, *_u_coeffs(m.mu, m.lam))
        return cache[key]

    calib = None
    if jump == "calibrated":
        calib = {}
        for region in model.regions:
            for q in region.patches:
                C = np.zeros((q.n_nodes, 3, 3))
                for p in region.patches:
                    blk = block(q, p, KERNEL_T, region)
                    C -= float(model.orientation(region, p)) * blk.reshape(
                        q.n_nodes, 3, p.n_nodes, 3).sum(axis=2)
                calib[(id(region), id(q))] = C
    n = layout.n_unknowns
    A = np.zeros((n, n))
    for t in system.terms:
        blk = block(t.field_patch, t.source_patch, t.kernel, t.region)
        r0, r1, c0, c1 = t.row.offset, t.row.stop, t.col.offset, t.col.stop
        A[r0:r1, c0:c1] += t.scale * blk
        if t.diag:
            D = calib[(id(t.region), id(t.field_patch))] if calib is not None else \
                np.broadcast_to(t.diag * np.eye(3), (t.field_patch.n_nodes, 3, 3))
            c = np.arange(D.shape[0])
            for a in range(3):
                for bb in range(3):
                    A[r0 + 3 * c + a, c0 + 3 * c + bb] += D[:, a, bb]
    b = np.zeros(n)
    for rt in system.rhs_terms:
        blk = block(rt.field_patch, rt.source_patch, rt.kernel, rt.region)
        b[rt.row.offset:rt.row.stop] += rt.scale * (blk @ rt.vector)
    for rv in system.rhs_values:
        b[rv.row.offset:rv.row.stop] += rv.vector
    for rt in system.rhs_terms:
        if not rt.add_half_of_vector:
            continue
        u_c = np.asarray(rt.vector, float).reshape(-1, 3)
        if calib is None:
            b[rt.row.offset:rt.row.stop] += rt.scale_half * u_c.ravel()
        else:
            D = calib[(id(rt.region), id(rt.field_patch))]
            b[rt.row.offset:rt.row.stop] -= np.einsum("nij,nj->ni", D, u_c).ravel()
    return A, b, calib


def lu_solve(A, b):
    """LU with partial pivoting (la/lu.py:110-161, "double")."""
    return sla.lu_solve(sla.lu_factor(A), b)


def split(layout, x):
    return {s.name: x[s.offset:s.stop].reshape(-1, 3) for s in layout.slots}


def fgmres(A_mv, b, M=None, x0=None, rtol=1e-8, restart=200, maxiter=600,
           stagnation_window=100, stagnation_factor=10.0):
    """la/solver.py:115-245 (modified Gram-Schmidt)."""
    M_mv = M if M is not None else (lambda v: v)
    b = np.asarray(b)
    n = b.shape[0]
    b_norm = np.linalg.norm(b)
    if b_norm == 0.0:
        return np.zeros(n), {"converged": True, "iterations": 0, "true_relres": 0.0}
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=float)
    history = []
    total_iters = 0
    best_res, best_res_at, stagnated = np.inf, 0, False
    while total_iters < maxiter and not stagnated:
        r = b - A_mv(x)
        beta = np.linalg.norm(r)
        if beta / b_norm < rtol:
            break
        m = min(restart, maxiter - total_iters)
        V = np.empty((m + 1, n))
        Z = np.empty((m, n))
        H = np.zeros((m + 1, m))
        cs, sn, g = np.zeros(m), np.zeros(m), np.zeros(m + 1)
        V[0] = r / beta
        g[0] = beta
        j_done = 0
        for j in range(m):
            Z[j] = M_mv(V[j])
            w = A_mv(Z[j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[i], w)
                w -= H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 1e-300:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            denom = np.sqrt(np.abs(H[j, j]) ** 2 + np.abs(H[j + 1, j]) ** 2)
            if denom == 0.0:
                j_done = j + 1
                break
            cs[j] = np.abs(H[j, j]) / denom if np.abs(H[j, j]) > 0 else 0.0
            if np.abs(H[j, j]) > 0:
                sn[j] = (H[j, j] / np.abs(H[j, j])) * H[j + 1, j] / denom
            else:
                sn[j] = 1.0
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total_iters += 1
            j_done = j + 1
            res = float(np.abs(g[j + 1])) / b_norm
            history.append(res)
            if res < best_res / stagnation_factor:
                best_res, best_res_at = res, total_iters
            elif total_iters - best_res_at >= stagnation_window:
                stagnated = True
                break
            if res < rtol or total_iters >= maxiter:
                break
        if j_done > 0:
            R = H[:j_done, :j_done]
            y = np.zeros(j_done)
            for i in range(j_done - 1, -1, -1):
                y[i] = (g[i] - R[i, i + 1:] @ y[i + 1:]) / R[i, i]
            x = x + Z[:j_done].T @ y
        if stagnated:
            break
    true_relres = float(np.linalg.norm(b - A_mv(x)) / b_norm)
    return x, {"converged": true_relres < rtol, "iterations": total_iters,
               "true_relres": true_relres, "history": history}


class BlockJacobiGS:
    """la/preconditioner.py:314-543 on the CPU: patch super-blocks in slot
    order, each either whole (dense LU) or cut into cluster chunks of at most
    ``bj_chunk`` DOFs; one forward Gauss-Seidel sweep."""

    def __init__(self, A, layout, points_of, max_dense=0, bj_chunk=9000):
        from mbem_amd.la.cluster import build_cluster_tree, leaves
        self.A = A
        self.blocks = []
        for slot in layout.slots:
            pts = points_of(slot.patch)
            if slot.size <= max_dense:
                parts = [np.arange(pts.shape[0])]
            else:
                parts = leaves(build_cluster_tree(pts, max(32, bj_chunk // 3)))
            chunks = []
            for p in parts:
                p = np.asarray(p)
                dof = (slot.offset + 3 * p[:, None] + np.arange(3)[None, :]).ravel()
                chunks.append((dof, sla.lu_factor(A[np.ix_(dof, dof)])))
            self.blocks.append((slot.offset, slot.stop, chunks))

    def __call__(self, r):
        z = np.zeros_like(r)
        for r0, r1, chunks in self.blocks:
            rk = r[r0:r1].copy()
            if r0 > 0:
                rk -= self.A[r0:r1, :r0] @ z[:r0]
            full = np.zeros_like(r)
            full[r0:r1] = rk
            for dof, lu in chunks:
                z[dof] = sla.lu_solve(lu, full[dof])
        return z


def _gcro_lstsq(G, rhs):
    y, *_ = np.linalg.lstsq(G, rhs, rcond=None)
    return y, float(np.linalg.norm(rhs - G @ y))


def _gcro_G(d, B, Hbar):
    """[[diag(d), B], [0, Hbar]] (la/solver.py:398-407)."""
    k = 0 if d is None else len(d)
    G = np.zeros((k + Hbar.shape[0], k + Hbar.shape[1]))
    if k:
        G[:k, :k] = np.diag(d)
        G[:k, k:] = B
    G[k:, k:] = Hbar
    return G


def harmonic_ritz_span(G, WtV, keep):
    """Real orthonormal basis of the k smallest-|theta| eigenvectors of the
    pencil G^T G z = theta G^T WtV z (la/solver.py:464-492)."""
    import scipy.linalg as sl
    theta, vec = sl.eig(G.T @ G, G.T @ WtV)
    mag = np.where(np.isfinite(np.abs(theta)), np.abs(theta), np.inf)
    idx = np.argsort(mag)[:keep]
    if idx.size == 0 or not np.isfinite(mag[idx[0]]):
        return None
    sel = vec[:, idx]
    parts = [sel.real] + ([sel.imag] if np.abs(sel.imag).max() > 0 else [])
    Q, s, _ = np.linalg.svd(np.hstack(parts), full_matrices=False)
    rank = int((s > 1e-12 * s[0]).sum()) if s.size else 0
    if rank == 0:
        return None
    return Q[:, :min(rank, keep + 1, G.shape[1])]


def gcrodr(A_mv, b, M=None, x0=None, Y=None, dim=30, rtol=1e-8, restart=200, maxiter=600,
           stagnation_window=100, stagnation_factor=10.0):
    """la/solver.py:248-395: GCRO-DR, right-preconditioned and flexible,
    modified Gram-Schmidt Arnoldi.  ``Y`` (n x k) is the recycled space in
    the A M domain; returns (x, info) with info["Y"] the space to carry."""
    import scipy.linalg as sl
    M_mv = M if M is not None else (lambda v: v)
    b = np.asarray(b, dtype=float)
    n = b.shape[0]
    bn = np.linalg.norm(b)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=float)
    C = U = None
    setup = 0
    if Y is not None and Y.ndim == 2 and Y.shape[0] == n and Y.shape[1] > 0:
        U0 = np.column_stack([M_mv(Y[:, i]) for i in range(Y.shape[1])])
        AU = np.column_stack([A_mv(U0[:, i]) for i in range(Y.shape[1])])
        setup = Y.shape[1]
        Q, R = np.linalg.qr(AU)
        dg = np.abs(np.diag(R))
        if dg.min() > 1e-12 * max(dg.max(), 1e-300):
            C = Q
            Rinv = sl.solve_triangular(R, np.eye(R.shape[0]))
            U, Y = U0 @ Rinv, Y @ Rinv
        else:
            Y = None
    else:
        Y = None
    r = b - A_mv(x)
    if C is not None:
        c = C.T @ r
        x = x + U @ c
        r = r - C @ c
    hist, total, cycles = [], 0, 0
    best, best_at, stag = np.inf, 0, False
    res = np.linalg.norm(r) / bn
    while res >= rtol and total < maxiter and not stag:
        k = 0 if C is None else C.shape[1]
        m = min(restart - k, maxiter - total)
        if m < 1:
            break
        beta = np.linalg.norm(r)
        V = np.zeros((n, m + 1))
        Z = np.zeros((n, m))
        Hb = np.zeros((m + 1, m))
        B = np.zeros((k, m))
        V[:, 0] = r / beta
        d = 1.0 / np.linalg.norm(Y, axis=0) if k else None
        head = C.T @ r if k else np.zeros(0)
        jd = 0
        for j in range(m):
            Z[:, j] = M_mv(V[:, j])
            w = A_mv(Z[:, j])
            if k:
                B[:, j] = C.T @ w
                w = w - C @ B[:, j]
            for i in range(j + 1):
                Hb[i, j] = V[:, i] @ w
                w = w - Hb[i, j] * V[:, i]
            Hb[j + 1, j] = np.linalg.norm(w)
            V[:, j + 1] = w / Hb[j + 1, j] if Hb[j + 1, j] > 1e-300 else 0.0
            total += 1
            jd = j + 1
            G = _gcro_G(d, B[:, :jd], Hb[:jd + 1, :jd])
            _, ra = _gcro_lstsq(G, np.concatenate([head, [beta], np.zeros(jd)]))
            res = ra / bn
            hist.append(res)
            if res < best / stagnation_factor:
                best, best_at = res, total
            elif total - best_at >= stagnation_window:
                stag = True
                break
            if res < rtol or total >= maxiter:
                break
        if jd == 0:
            break
        G = _gcro_G(d, B[:, :jd], Hb[:jd + 1, :jd])
        rhs = np.concatenate([head, [beta], np.zeros(jd)])
        y, _ = _gcro_lstsq(G, rhs)
        x = x + Z[:, :jd] @ y[k:]
        if k:
            x = x + (U * d) @ y[:k]
        g = rhs - G @ y
        r = V[:, :jd + 1] @ g[k:] + (C @ g[:k] if k else 0.0)
        cycles += 1
        if stag:
            break
        # the next space: harmonic Ritz vectors of the augmented pencil
        Yh = Y * d if k else None
        Uh = U * d if k else None
        W = np.zeros((k + jd + 1, k + jd))
        W[k:k + jd, k:] = np.eye(jd)
        if k:
            W[:k, :k] = C.T @ Yh
            W[k:, :k] = V[:, :jd + 1].T @ Yh
        P = harmonic_ritz_span(G, W, min(dim, k + jd))
        if P is None:
            C = U = Y = None
            continue
        Yn = V[:, :jd] @ P[k:] + (Yh @ P[:k] if k else 0.0)
        Un = Z[:, :jd] @ P[k:] + (Uh @ P[:k] if k else 0.0)
        Q, R = np.linalg.qr(G @ P)
        dg = np.abs(np.diag(R))
        if dg.size == 0 or dg.min() <= 1e-12 * max(dg.max(), 1e-300):
            C = U = Y = None
            continue
        Cn = V[:, :jd + 1] @ Q[k:] + (C @ Q[:k] if k else 0.0)
        Rinv = sl.solve_triangular(R, np.eye(R.shape[0]))
        C, U, Y = Cn, Un @ Rinv, Yn @ Rinv
    tr = float(np.linalg.norm(b - A_mv(x)) / bn)
    Yout = None if Y is None else Y / np.linalg.norm(Y, axis=0)
    return x, {"converged": tr < rtol, "iterations": total, "true_relres": tr, "history": hist,
               "Y": Yout, "setup_matvecs": setup, "cycles": cycles}
