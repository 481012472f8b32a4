"""Flexible right-preconditioned GMRES on device-resident vectors
(mbem's src/mbem/la/solver.py:115-245, the plain path).

Same contract as the reference: ``fgmres(A, b, M, x0, rtol, restart, maxiter,
stagnation_window, stagnation_factor, callback) -> (x, SolveReport)``, right
preconditioning, Givens rotations in the reference's ``phase`` form, the
``H[j+1, j] > 1e-300`` breakdown guard, Arnoldi residual ``|g[j+1]| / ||b||``,
stagnation when no factor-``stagnation_factor`` drop in
``stagnation_window`` iterations, ``x += Z^T y`` at each restart and the TRUE
residual ``||b - A x|| / ||b||`` deciding ``converged``; ``||b|| = 0`` gives
zeros.

Where it differs, and why: the Krylov basis V (restart+1, n) and the
flexible basis Z (restart, n) live in HBM and the orthogonalization is
classical Gram-Schmidt applied twice (CGS2: two multi-dot + multi-axpy
passes, each one launch over the whole basis) instead of modified
Gram-Schmidt's j+1 dependent dot/axpy pairs -- the same projection in exact
arithmetic and as stable as MGS in floating point, at two HBM passes per
iteration.  The small Hessenberg least-squares problem stays on the host in
FP64, exactly the reference's update order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib, defaults
from .._lib import call, ptr, stream

F64 = torch.float64


@dataclass
class SolveReport:
    """solver.py:31-68."""
    converged: bool
    iterations: int
    true_relres: float
    arnoldi_relres: float
    stagnated: bool = False
    restarts: int = 0
    residual_history: list = field(default_factory=list)
    precond_summary: dict | None = None
    recycle_dim: int = 0
    setup_matvecs: int = 0
    matvecs: int = 0
    precond_applies: int = 0

    def __str__(self):
        status = "converged" if self.converged else (
            "STAGNATED" if self.stagnated else "NOT CONVERGED")
        s = (f"FGMRES {status}: {self.iterations} iters, "
             f"true relres {self.true_relres:.3e}")
        if self.precond_summary:
            counts: dict = {}
            for sb in self.precond_summary["super_blocks"]:
                counts[sb["rung"]] = counts.get(sb["rung"], 0) + 1
            rungs = ", ".join(f"{n} {r}" for r, n in counts.items())
            s += f"; preconditioner {rungs}, built in {self.precond_summary['build_s']:.2f} s"
        return s


class RecycleSpace:
    """The subspace GCRO-DR carries from one solve of a sequence to the next
    (solver.py:72-112): ``vectors`` (n x k') in the domain of the
    preconditioned operator A M, with unit columns; each solve re-derives
    U = M Y and C = A U for its own A and M, so both may change between
    solves.  Held on the device as k' rows; ``reset()`` drops them (a changed
    layout or a material step past the reuse window)."""

    def __init__(self, dim: int = defaults.GCRO_RECYCLE_DIM):
        self.dim = int(dim)
        self._rows = None

    @property
    def vectors(self):
        return None if self._rows is None else self._rows.T

    @vectors.setter
    def vectors(self, Y):
        if Y is None:
            self._rows = None
            return
        Y = torch.as_tensor(Y, dtype=F64, device=_lib.device())
        self._store(Y.T.contiguous())

    def reset(self):
        self._rows = None

    def _store(self, rows):
        """Keep the span with unit rows (zero rows dropped), as a copy."""
        if rows is None or rows.numel() == 0:
            self._rows = None
            return
        norms = torch.linalg.vector_norm(rows, dim=1)
        keep = norms > 0.0
        self._rows = (rows[keep] / norms[keep, None]).contiguous() if bool(keep.any()) else None


def _solve_upper(R, g):
    """Back substitution of the small upper-triangular system (solver.py:503-510)."""
    m = R.shape[0]
    y = np.zeros(m, dtype=R.dtype)
    for i in range(m - 1, -1, -1):
        s = g[i] - R[i, i + 1:] @ y[i + 1:]
        y[i] = s / R[i, i]
    return y


def _as_op(A, host: bool = False):
    """Operator ``op(v, out)`` writing A v into ``out`` (solver.py:513-520):
    a device matrix (or a host ndarray, uploaded once) is applied with the
    library's GEMV, an object with ``apply_into(v, out)`` writes in place,
    and ``.matvec`` objects, plain callables and objects with ``@`` return a
    new vector that is copied in.  With ``host`` (the caller handed a numpy
    right-hand side: the reference's own callers) those callables see numpy
    vectors and may return numpy or torch."""
    if A is None:
        return None
    if isinstance(A, np.ndarray):
        A = torch.as_tensor(np.ascontiguousarray(A, dtype=np.float64), device=_lib.device())
    if isinstance(A, torch.Tensor):
        A = A.to(device=_lib.device(), dtype=F64).contiguous()
        return lambda v, out: gemv(A, v, out)
    if hasattr(A, "apply_into"):
        return A.apply_into
    if callable(A) and not hasattr(A, "matvec"):
        f = A
    elif hasattr(A, "matvec"):
        f = A.matvec
    else:
        f = lambda v: A @ v  # noqa: E731
    if not host:
        return lambda v, out: out.copy_(f(v))

    def op(v, out):
        y = f(v.cpu().numpy())
        out.copy_(torch.as_tensor(np.asarray(y, dtype=np.float64)) if not
                  isinstance(y, torch.Tensor) else y)
        return out
    return op


def gemv(A: torch.Tensor, x: torch.Tensor, out=None, alpha=1.0, beta=0.0):
    """out = alpha A x + beta out on the device (mbem_gemv)."""
    m, n = A.shape
    if out is None:
        out = torch.empty(m, dtype=F64, device=A.device)
        beta = 0.0
    call("mbem_gemv", ptr(A), A.stride(0), m, n, ptr(x), ptr(out), float(alpha), float(beta),
         stream())
    return out


class _Krylov:
    """Device workspace and the three vector passes of one FGMRES."""

    def __init__(self, n, m, dev, tail: int = 0):
        self.n = n
        self.V = torch.empty((m + 1, n), dtype=F64, device=dev)
        self.Z = torch.empty((m, n), dtype=F64, device=dev)
        # CGS2 coefficients and norm in h[:2 (m + 1) + 1]; ``tail`` more
        # entries behind them (``tail_view``) travel to the host with them
        self.h_len = 2 * (m + 1) + 1
        self.h = torch.empty(self.h_len + tail, dtype=F64, device=dev)
        self.work = torch.empty(int(_lib.load().mbem_multidot_work(m + 1, n)), dtype=F64,
                                device=dev)
        self._h_host = None          # pinned copy of h (cgs2 with before_wait)
        self._h_ready = None

    def dots(self, rows: torch.Tensor, k: int, w: torch.Tensor, at: int = 0) -> torch.Tensor:
        """h[at:at+k] = rows[:k] . w on the device (no host sync)."""
        out = self.h[at:at + k]
        call("mbem_multidot", ptr(rows), rows.stride(0), k, self.n, ptr(w), ptr(out),
             ptr(self.work), stream())
        return out

    def tail_view(self, e: int) -> torch.Tensor:
        """The first e entries of h's tail (a device buffer whose values
        ``cgs2(..., tail=e)`` returns with its own)."""
        return self.h[self.h_len:self.h_len + e]

    def cgs2(self, V: torch.Tensor, k: int, w: torch.Tensor, normalize: bool = False,
             tail: int = 0, before_wait=None):
        """Project w against V[:k] twice (classical Gram-Schmidt, reorthogonalized)
        and return (h, ||w||) on the host after ONE synchronization; with
        ``normalize`` w is also divided by its norm on the device (when the
        norm exceeds 1e-300) before that synchronization.  With ``tail`` = e
        the first e values of ``tail_view`` come back in the same transfer,
        as a third element.  ``before_wait`` (a callable) is called once the
        transfer is queued and before the host waits for it: work it queues
        runs on the device while the host waits (the transfer waits only for
        the work queued before it)."""
        h1 = self.dots(V, k, w, 0)
        self.axpy_rows(V, k, h1, w, -1.0)
        h2 = self.dots(V, k, w, k)
        self.axpy_rows(V, k, h2, w, -1.0)
        self.dots(w.view(1, -1), 1, w, 2 * k)
        if normalize:
            call("mbem_normalize", self.n, ptr(self.h[2 * k:]), ptr(w), stream())
        e = self.h_len + tail if tail else 2 * k + 1
        if before_wait is None:
            hv = self.h[:e].cpu().numpy()
        else:
            if self._h_host is None:
                self._h_host = torch.empty(self.h.numel(), dtype=F64, pin_memory=True)
                self._h_ready = torch.cuda.Event()
            self._h_host[:e].copy_(self.h[:e], non_blocking=True)
            self._h_ready.record()
            before_wait()
            self._h_ready.synchronize()
            hv = self._h_host[:e].numpy().copy()
        if not tail:
            return hv[:k] + hv[k:2 * k], float(np.sqrt(max(hv[2 * k], 0.0)))
        return (hv[:k] + hv[k:2 * k], float(np.sqrt(max(hv[2 * k], 0.0))),
                hv[self.h_len:self.h_len + tail].copy())

    def axpy_rows(self, rows: torch.Tensor, k: int, coef: torch.Tensor, w: torch.Tensor,
                  alpha: float):
        call("mbem_multiaxpy", ptr(rows), rows.stride(0), k, self.n, ptr(coef), ptr(w),
             float(alpha), stream())

    def norm(self, w: torch.Tensor) -> float:
        d = self.dots(w.view(1, -1), 1, w)
        return float(np.sqrt(max(float(d.item()), 0.0)))


def axpby(a, x, b, y):
    call("mbem_axpby", y.numel(), float(a), ptr(x), float(b), ptr(y), stream())
    return y


def fgmres(A, b, M=None, x0=None,
           rtol: float = defaults.GMRES_RTOL,
           restart: int = defaults.GMRES_RESTART,
           maxiter: int = defaults.GMRES_MAXITER,
           stagnation_window: int = defaults.GMRES_STAGNATION_WINDOW,
           stagnation_factor: float = defaults.GMRES_STAGNATION_FACTOR,
           callback=None, recycle=None):
    """Solve A x = b on the device; returns (x, SolveReport).

    ``A`` and ``M`` are matrices (device tensors or host ndarrays), objects
    with ``apply_into`` / ``matvec`` / ``@``, or callables ``f(v) -> A v``;
    ``M`` approximates A^-1 (right preconditioner).  A torch ``b`` gives a
    device ``x`` and callables see device vectors; a numpy ``b`` (the
    reference's callers) gives a numpy ``x`` and callables see numpy
    vectors.  A complex ``b`` raises TypeError, as in the reference."""
    _lib.require_gpu()
    host = not isinstance(b, torch.Tensor)
    if host:
        b = np.asarray(b)
        if np.iscomplexobj(b):
            raise TypeError("fgmres is real-only")
        b = torch.as_tensor(np.asarray(b, dtype=np.float64))
    elif b.is_complex():
        raise TypeError("fgmres is real-only")
    if host:
        x, rep = _fgmres_device(_as_op(A, True), _as_op(M, True) if M is not None else None, b,
                                x0, rtol, restart, maxiter, stagnation_window, stagnation_factor,
                                callback, recycle)
        return x.cpu().numpy(), rep
    return _fgmres_device(_as_op(A), _as_op(M) if M is not None else None, b, x0, rtol,
                          restart, maxiter, stagnation_window, stagnation_factor, callback,
                          recycle)


def _fgmres_device(A_mv, M_mv, b, x0, rtol, restart, maxiter, stagnation_window,
                   stagnation_factor, callback, recycle):
    """The iteration of ``fgmres`` on device vectors, with ``A_mv`` / ``M_mv``
    already in the ``op(v, out)`` form."""
    b = b.to(device=_lib.device(), dtype=F64).contiguous()
    n = b.shape[0]
    dev = b.device
    ws = _Krylov(n, max(1, min(restart, maxiter)), dev)
    b_norm = ws.norm(b)
    if b_norm == 0.0:
        return torch.zeros(n, dtype=F64, device=dev), SolveReport(True, 0, 0.0, 0.0)
    if x0 is not None and not isinstance(x0, torch.Tensor):
        x0 = torch.as_tensor(np.asarray(x0, dtype=np.float64))
    x = torch.zeros(n, dtype=F64, device=dev) if x0 is None else \
        x0.to(device=dev, dtype=F64).clone().contiguous()
    if recycle is not None:
        # GCRO-DR with the subspace carried in ``recycle`` (solver.py:248-395)
        from .gcrodr import GcroDR
        g = GcroDR(n, max(2, restart), recycle.dim, dev)
        return g.solve(A_mv, M_mv if M_mv is not None else (lambda v, out: out.copy_(v)), b, x,
                       x0 is None, recycle, rtol, maxiter, stagnation_window, stagnation_factor,
                       callback)

    history: list = []
    total_iters = 0
    restarts = 0
    best_res = np.inf
    best_res_at = 0
    stagnated = False
    n_mv = 0
    n_pc = 0
    r = torch.empty(n, dtype=F64, device=dev)

    x_is_zero = x0 is None
    r_is_current = False                             # r holds b - A x for the current x
    while total_iters < maxiter and not stagnated:
        if x_is_zero:
            r.copy_(b)                               # r = b - A 0, no matvec
            x_is_zero = False
        else:
            A_mv(x, r)
            n_mv += 1
            axpby(1.0, b, -1.0, r)                   # r = b - A x
        r_is_current = True
        beta = ws.norm(r)
        arnoldi_res = beta / b_norm
        if arnoldi_res < rtol:
            break
        m = min(restart, maxiter - total_iters)
        V, Z = ws.V, ws.Z
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        axpby(1.0 / beta, r, 0.0, V[0])
        g[0] = beta
        j_done = 0
        for j in range(m):
            if M_mv is None:
                Z[j].copy_(V[j])
            else:
                M_mv(V[j], Z[j])
                n_pc += 1
            w = V[j + 1]
            A_mv(Z[j], w)
            n_mv += 1
            # classical Gram-Schmidt, twice, and V[j+1] = w / ||w|| (when
            # ||w|| > 1e-300) on the device; one host synchronization
            hsum, wnorm = ws.cgs2(V, j + 1, w, normalize=True)
            H[:j + 1, j] = hsum
            H[j + 1, j] = wnorm
            # Givens (solver.py:187-205)
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
                phase = H[j, j] / np.abs(H[j, j])
                sn[j] = phase * H[j + 1, j] / denom
            else:
                sn[j] = 1.0
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total_iters += 1
            j_done = j + 1
            arnoldi_res = float(np.abs(g[j + 1])) / b_norm
            history.append(arnoldi_res)
            if callback is not None:
                callback(total_iters, arnoldi_res)
            if arnoldi_res < best_res / stagnation_factor:
                best_res = arnoldi_res
                best_res_at = total_iters
            elif total_iters - best_res_at >= stagnation_window:
                stagnated = True
                break
            if arnoldi_res < rtol or total_iters >= maxiter:
                break
        if j_done > 0:
            y = _solve_upper(H[:j_done, :j_done], g[:j_done])
            yd = torch.as_tensor(y, dtype=F64, device=dev)
            ws.axpy_rows(Z, j_done, yd, x, 1.0)
            r_is_current = False
        restarts += 1
        if stagnated:
            break

    # the true residual of the returned x; when the restart loop ended right
    # after computing it (a converged restart check), that r is reused
    # instead of a second product with A
    if not r_is_current:
        A_mv(x, r)
        n_mv += 1
        axpby(1.0, b, -1.0, r)
    true_relres = ws.norm(r) / b_norm
    arnoldi_final = history[-1] if history else true_relres
    converged = true_relres < rtol
    return x, SolveReport(converged=converged, iterations=total_iters, true_relres=true_relres,
                          arnoldi_relres=float(arnoldi_final),
                          stagnated=stagnated and not converged, restarts=restarts - 1,
                          residual_history=history, matvecs=n_mv, precond_applies=n_pc)
