"""GCRO with deflated restarting, right-preconditioned and flexible, on
device-resident vectors (mbem's src/mbem/la/solver.py:248-500, after Parks,
de Sturler, Mackey, Johnson & Maiti, SIAM J. Sci. Comput. 28(5) 1651-1674,
2006, applied to the right-preconditioned operator A M).

Same contract as the reference's ``fgmres(..., recycle=rs)``: the recycled
vectors Y live in the domain of A M; each solve re-derives U = M Y and
C = A U for its own A and M (k preconditioner applications and k matvecs
before the first iteration), orthonormalizes C (2.7) and takes the optimal
correction over range(U); cycles of m - k Arnoldi steps on (I - C C^T) A M
(2.9); the augmented least squares (2.11, 2.13) gives the per-iteration
residual estimate that stops the iteration and the cycle's update; the new
recycle space is spanned by the k harmonic Ritz vectors of smallest
magnitude of the pencil (2.16); the TRUE residual decides ``converged``.

Layout: every basis (V, Z, C, U, Y) is a row-major (count, n) device
matrix, so each product with a basis is one multi-dot or multi-axpy launch
over the whole basis; the small dense problems (least squares, QR of the
k x k factors, the generalized eigenproblem) stay on the host in FP64.
Where it differs from the reference: the Arnoldi vectors are orthogonalized
by classical Gram-Schmidt applied twice (as in the plain FGMRES, one host
synchronization per iteration), and the QR of A U by the same CGS2 column
by column (R with a positive diagonal; the span and A U = C are what
matter).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import torch

from .. import _lib
from .._lib import call, ptr, stream

F64 = torch.float64


class _Bases:
    """The (C, U, Y) triple: k rows each in preallocated (kmax, n) buffers."""

    def __init__(self, kmax, n, dev):
        self.C = torch.zeros((kmax, n), dtype=F64, device=dev)
        self.U = torch.zeros((kmax, n), dtype=F64, device=dev)
        self.Y = torch.zeros((kmax, n), dtype=F64, device=dev)
        self.k = 0


class _Ops:
    """Row-basis products on the device: dots, combinations, norms."""

    def __init__(self, n, rows_max, dev):
        self.n = n
        self.dev = dev
        self.work = torch.empty(int(_lib.load().mbem_multidot_work(rows_max, n)), dtype=F64,
                                device=dev)
        self.h = torch.empty(rows_max, dtype=F64, device=dev)

    def dots(self, rows, k, w):
        """rows[:k] . w on the host (one synchronization)."""
        if k == 0:
            return np.zeros(0)
        call("mbem_multidot", ptr(rows), rows.stride(0), k, self.n, ptr(w), ptr(self.h),
             ptr(self.work), stream())
        return self.h[:k].cpu().numpy()

    def dots_dev(self, rows, k, w, out):
        call("mbem_multidot", ptr(rows), rows.stride(0), k, self.n, ptr(w), ptr(out),
             ptr(self.work), stream())
        return out

    def axpy(self, rows, k, coef, w, alpha=1.0):
        """w += alpha rows[:k]^T coef; coef a host vector or a device one."""
        if k == 0:
            return
        if not isinstance(coef, torch.Tensor):
            coef = torch.as_tensor(np.ascontiguousarray(coef, dtype=np.float64)).to(self.dev)
        call("mbem_multiaxpy", ptr(rows), rows.stride(0), k, self.n, ptr(coef), ptr(w),
             float(alpha), stream())

    def combine(self, terms, out, q):
        """out[i] = sum over (rows, k, coef) in terms of rows[:k]^T coef[:, i],
        i < q; each coef (k x q) is uploaded once."""
        ups = []
        for rows, k, coef in terms:
            if k:
                ups.append((rows, k, torch.as_tensor(np.ascontiguousarray(
                    np.asarray(coef, dtype=np.float64).T)).to(self.dev)))
        for i in range(q):
            out[i].zero_()
            for rows, k, cT in ups:
                call("mbem_multiaxpy", ptr(rows), rows.stride(0), k, self.n, ptr(cT[i]),
                     ptr(out[i]), 1.0, stream())

    def norm(self, w):
        call("mbem_multidot", ptr(w), self.n, 1, self.n, ptr(w), ptr(self.h), ptr(self.work),
             stream())
        return float(np.sqrt(max(float(self.h[0].item()), 0.0)))


def _augmented_G(d, B, Hbar):
    """G of (2.11): [[D_k, B_k], [0, Hbar]] (no D, B block when k = 0)."""
    k = 0 if d is None else d.size
    nr, nc = Hbar.shape
    G = np.zeros((k + nr, k + nc))
    if k:
        G[:k, :k] = np.diag(d)
        G[:k, k:] = B
    G[k:, k:] = Hbar
    return G


def _lstsq(G, rhs):
    """Minimizer of |rhs - G y| over the small augmented system and the
    residual norm it leaves."""
    y, *_ = np.linalg.lstsq(G, rhs, rcond=None)
    return y, float(np.linalg.norm(rhs - G @ y))


def _harmonic_basis(G, WtV, k_keep):
    """Orthonormal real basis (search-space coordinates) of the span of the
    k_keep harmonic Ritz vectors of smallest |theta| of the pencil
    G^T G z = theta G^T W^T V z (2.16); a complex pair contributes its real
    and imaginary parts, so up to k_keep + 1 columns.  None if degenerate."""
    try:
        theta, vec = sla.eig(G.T @ G, G.T @ WtV)
    except (sla.LinAlgError, ValueError):
        return None
    mag = np.abs(theta)
    mag[~np.isfinite(mag)] = np.inf
    order = np.argsort(mag)[:k_keep]
    if order.size == 0 or not np.isfinite(mag[order[0]]):
        return None
    sel = vec[:, order]
    cols = [sel.real] + ([sel.imag] if np.any(np.abs(sel.imag) > 0) else [])
    Q, s, _ = np.linalg.svd(np.column_stack(cols), full_matrices=False)
    rank = int(np.sum(s > s[0] * 1e-12)) if s.size else 0
    if rank == 0:
        return None
    return Q[:, :min(rank, k_keep + 1, G.shape[1])]


class GcroDR:
    """Workspace of one recycled solve sequence on n unknowns."""

    def __init__(self, n, restart, dim, dev):
        from .solver import _Krylov
        self.n, self.m, self.dim, self.dev = n, restart, dim, dev
        kmax = dim + 2
        # the C-projection coefficients of an Arnoldi step sit in the tail of
        # the CGS2 buffer: both reach the host in one transfer per step
        self.ws = _Krylov(n, restart, dev, tail=kmax)
        self.sets = [_Bases(kmax, n, dev), _Bases(kmax, n, dev)]
        self.T1 = torch.zeros((kmax, n), dtype=F64, device=dev)
        self.T2 = torch.zeros((kmax, n), dtype=F64, device=dev)
        self.ops = _Ops(n, max(restart + 1, kmax), dev)
        self.hb = self.ws.tail_view(kmax)

    # ---- (2.7): A U = C with C orthonormal --------------------------------
    def _orthonormalize(self, AU, k, src_u, src_y, dst):
        """dst.C = Q of AU[:k] (CGS2, column by column), dst.U, dst.Y =
        (src_u, src_y) R^-1; dst.k = 0 when A U is rank deficient."""
        ops = self.ops
        R = np.zeros((k, k))
        for i in range(k):
            q = dst.C[i]
            q.copy_(AU[i])
            for _ in range(2):
                h = ops.dots(dst.C, i, q)
                ops.axpy(dst.C, i, h, q, -1.0)
                R[:i, i] += h
            R[i, i] = ops.norm(q)
            if R[i, i] > 0:
                q.mul_(1.0 / R[i, i])
        diag = np.abs(np.diag(R))
        if diag.size == 0 or diag.min() <= 1e-12 * max(diag.max(), 1e-300):
            dst.k = 0
            return
        Rinv = sla.solve_triangular(R, np.eye(k))
        ops.combine([(src_u, k, Rinv)], dst.U, k)
        ops.combine([(src_y, k, Rinv)], dst.Y, k)
        dst.k = k

    def solve(self, A_mv, M_mv, b, x, x_is_zero, rs, rtol, maxiter, stagnation_window,
              stagnation_factor, callback):
        from .solver import SolveReport, axpby
        n, ops, ws = self.n, self.ops, self.ws
        b_norm = ops.norm(b)
        cur, nxt = self.sets
        setup_matvecs = 0
        Yin = rs._rows
        if Yin is not None and (Yin.ndim != 2 or Yin.shape[1] != n or Yin.shape[0] == 0
                                or Yin.shape[0] > self.dim + 1):
            rs.reset()
            Yin = None
        cur.k = 0
        if Yin is not None:
            kin = Yin.shape[0]
            self.T2[:kin].copy_(Yin)
            for i in range(kin):
                M_mv(self.T2[i], self.T1[i])                 # U_in = M Y_in
                A_mv(self.T1[i], nxt.C[i])                   # A U_in (held in nxt.C)
            setup_matvecs = kin
            AU = nxt.C[:kin].clone()
            self._orthonormalize(AU, kin, self.T1, self.T2, cur)
            del AU
        r = torch.empty(n, dtype=F64, device=self.dev)
        if x_is_zero:
            r.copy_(b)
        else:
            A_mv(x, r)
            axpby(1.0, b, -1.0, r)                           # r = b - A x
        if cur.k:
            y0 = ops.dots(cur.C, cur.k, r)
            ops.axpy(cur.U, cur.k, y0, x, 1.0)
            ops.axpy(cur.C, cur.k, y0, r, -1.0)
        history = []
        total, restarts = 0, 0
        best, best_at = np.inf, 0
        stagnated = False
        res = ops.norm(r) / b_norm
        V, Z = ws.V, ws.Z
        while res >= rtol and total < maxiter and not stagnated:
            k = cur.k
            m = min(self.m - k, maxiter - total)
            if m < 1:
                break
            beta = ops.norm(r)
            V[0].copy_(r).mul_(1.0 / beta)
            H = np.zeros((m + 1, m))
            Bk = np.zeros((k, m))
            if k:
                ynorm = np.array([ops.norm(cur.Y[i]) for i in range(k)])
                d = 1.0 / ynorm                              # D_k of (2.11)
                rhs_head = ops.dots(cur.C, k, r)
            else:
                d, rhs_head = None, np.zeros(0)
            j_done = 0
            for j in range(m):
                M_mv(V[j], Z[j])
                w = V[j + 1]
                A_mv(Z[j], w)
                if k:                                        # (I - C C^T) A M
                    ops.dots_dev(cur.C, k, w, self.hb)
                    ops.axpy(cur.C, k, self.hb[:k], w, -1.0)
                    hsum, wnorm, bk = ws.cgs2(V, j + 1, w, normalize=True, tail=k)
                    Bk[:, j] = bk
                else:
                    hsum, wnorm = ws.cgs2(V, j + 1, w, normalize=True)
                H[:j + 1, j] = hsum
                H[j + 1, j] = wnorm
                if not wnorm > 1e-300:
                    w.zero_()                                # lucky breakdown
                total += 1
                j_done = j + 1
                G = _augmented_G(d, Bk[:, :j_done], H[:j_done + 1, :j_done])
                rhs = np.concatenate([rhs_head, [beta], np.zeros(j_done)])
                _, res_abs = _lstsq(G, rhs)
                res = res_abs / b_norm
                history.append(res)
                if callback is not None:
                    callback(total, res)
                if res < best / stagnation_factor:
                    best, best_at = res, total
                elif total - best_at >= stagnation_window:
                    stagnated = True
                    break
                if res < rtol or total >= maxiter:
                    break
            if j_done == 0:
                break
            # (2.13-2.15): solution update and the cycle's residual
            G = _augmented_G(d, Bk[:, :j_done], H[:j_done + 1, :j_done])
            rhs = np.concatenate([rhs_head, [beta], np.zeros(j_done)])
            y, _ = _lstsq(G, rhs)
            ops.axpy(Z, j_done, y[k:], x, 1.0)
            if k:
                ops.axpy(cur.U, k, d * y[:k], x, 1.0)         # U D y
            g = rhs - G @ y
            r.zero_()
            ops.axpy(V, j_done + 1, g[k:], r, 1.0)
            if k:
                ops.axpy(cur.C, k, g[:k], r, 1.0)
            restarts += 1
            if stagnated:
                break
            self._recycle_update(cur, nxt, d, j_done, G, rs.dim)
            cur, nxt = nxt, cur
        rs._store(cur.Y[:cur.k] if cur.k else None)
        A_mv(x, r)
        axpby(1.0, b, -1.0, r)
        true_relres = ops.norm(r) / b_norm
        return x, SolveReport(
            converged=true_relres < rtol, iterations=total, true_relres=true_relres,
            arnoldi_relres=float(history[-1] if history else true_relres),
            stagnated=stagnated and not true_relres < rtol, restarts=max(restarts - 1, 0),
            residual_history=history, recycle_dim=cur.k, setup_matvecs=setup_matvecs)

    # ---- (2.16, A.30-A.34): the space carried to the next cycle ----------
    def _recycle_update(self, cur, nxt, d, j, G, k_keep):
        ops = self.ops
        V, Z = self.ws.V, self.ws.Z
        k = cur.k
        nv = j + 1
        WtV = np.zeros((k + nv, k + j))
        WtV[k:k + j, k:] = np.eye(j)
        for b_ in range(k):
            # columns of C^T Yh and V^T Yh, Yh = Y D
            WtV[:k, b_] = d[b_] * ops.dots(cur.C, k, cur.Y[b_])
            WtV[k:, b_] = d[b_] * ops.dots(V, nv, cur.Y[b_])
        P = _harmonic_basis(G, WtV, min(k_keep, k + j))
        if P is None:
            nxt.k = 0
            return
        q = P.shape[1]
        if q > nxt.C.shape[0]:
            P = P[:, :nxt.C.shape[0]]
            q = P.shape[1]
        Pd = P[:k] * d[:, None] if k else None              # Yh P = Y (D P)
        # Y_new = V P[k:] + Yh P[:k], U_new = Z P[k:] + Uh P[:k]
        self.ops.combine([(V, j, P[k:])] + ([(cur.Y, k, Pd)] if k else []), self.T2, q)
        self.ops.combine([(Z, j, P[k:])] + ([(cur.U, k, Pd)] if k else []), self.T1, q)
        Q, R = np.linalg.qr(G @ P)
        diag = np.abs(np.diag(R))
        if diag.size == 0 or diag.min() <= 1e-12 * max(diag.max(), 1e-300):
            nxt.k = 0
            return
        # C_new = V Q[k:] + C Q[:k]; (U, Y)_new scaled by R^-1
        self.ops.combine([(V, nv, Q[k:])] + ([(cur.C, k, Q[:k])] if k else []), nxt.C, q)
        Rinv = sla.solve_triangular(R, np.eye(q))
        self.ops.combine([(self.T1, q, Rinv)], nxt.U, q)
        self.ops.combine([(self.T2, q, Rinv)], nxt.Y, q)
        nxt.k = q
