"""Mixed-precision solve of an assembled dense system.

Operator split A = F + N on the device:
  N  near-field and self-element 3x3 blocks in FP64 block-CSR -- every
     (collocation point, source triangle) pair closer than
     ``k_near * (eps + h)`` of the source, where the closed-form log and
     arctangent terms cancel;
  F  everything else in BF16, FP32 or int8 storage, near entries zeroed,
     multiplied with FP32 accumulation (int8: offset-binary bytes with one
     FP32 scale per row and 64-column segment).

Solve: the outer loop is FP64 FGMRES on the FP64 operator (Arnoldi, Givens
and the true residual all FP64), so it is iterative refinement -- either the
assembled FP64 matrix or, with int8 storage, ``SplitOuter``: the int8 codes
plus their 24-bit rounding remainders, every far entry within 2e-10 of its
segment's largest magnitude, from 4 bytes instead of 8; its
right preconditioner is an inner FGMRES on the mixed operator F + N,
preconditioned by block Gauss-Seidel / block-Jacobi with FP32 chunk
inverses and mixed-operator couplings, run to a loose tolerance.  The error
against the FP64 solution is set by the outer rtol (the split operator's
entry errors, about 2e-10 of each segment's largest, add 0.8e-9 to 4.5e-9
at n = 60,000).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib, defaults
from .._lib import call, ptr, stream
from ..kernels import basis as kb
from .solver import SolveReport, _Krylov, _solve_upper, axpby, fgmres

F64 = torch.float64
I64 = torch.int64


def near_inputs(assembled, k_near: float = 2.0):
    """(pts, cen, rad) per unknown block: collocation point of row block i,
    source centroid of column block j and its near radius k_near (eps + h)."""
    layout = assembled.layout
    n_blk = layout.n_unknowns // 3
    pts = np.empty((n_blk, 3))
    cen = np.empty((n_blk, 3))
    rad = np.empty(n_blk)
    for s in layout.slots:
        p = s.patch
        b0, b1 = s.offset // 3, s.stop // 3
        pts[b0:b1] = assembled.collocation_points_of(p)
        cen[b0:b1] = p.mesh.centroids()
        rad[b0:b1] = k_near * (assembled.eps_for(p) + kb.element_sizes(p.mesh))
    return pts, cen, rad


class NearPattern:
    """The near-field block pattern in two phases: the constructor uploads
    the geometry (pinned, non-blocking) and queues the counting pass on the
    current stream; ``result()`` reads the counts, queues the filling pass
    and returns (rowptr, col) as host arrays.  A caller can do other host
    work between the two while the counts are computed."""

    def __init__(self, assembled, k_near: float = 2.0):
        pts, cen, rad = near_inputs(assembled, k_near)
        self.nb = pts.shape[0]
        self.P, self.C, self.Rd = (_lib.upload(a, F64) for a in (pts, cen, rad))
        self.stream = torch.cuda.current_stream()
        self.counts = torch.empty(self.nb, dtype=I64, device=self.P.device)
        call("mbem_near_pattern", ptr(self.P), ptr(self.C), ptr(self.Rd), self.nb,
             ptr(self.counts), None, None, stream())
        self._out = None

    def result(self) -> tuple:
        if self._out is None:
            with torch.cuda.stream(self.stream):
                nb, dev = self.nb, self.P.device
                rowptr = torch.zeros(nb + 1, dtype=I64, device=dev)
                rowptr[1:] = torch.cumsum(self.counts, 0)
                nnz = int(rowptr[-1].item())
                col = torch.empty(max(nnz, 1), dtype=I64, device=dev)
                call("mbem_near_pattern", ptr(self.P), ptr(self.C), ptr(self.Rd), nb, None,
                     ptr(rowptr), ptr(col), stream())
                self._out = (rowptr.cpu().numpy(), col[:nnz].cpu().numpy())
            self.P = self.C = self.Rd = self.counts = None
        return self._out


def near_structure(assembled, k_near: float = 2.0):
    """Block-CSR pattern (rowptr, col) of the near field over the system's
    unknown blocks (row block i = collocation point of unknown triple i,
    column block j = source element of unknown triple j): near when
    |x_i - c_j| < k_near * (eps_j + h_j).  Counted and filled on the GPU
    (mbem_near_pattern); returned as host arrays."""
    return NearPattern(assembled, k_near).result()


class MixedOperator:
    """y = F x + N x on the device (see the module docstring)."""

    def __init__(self, assembled, storage: str = "bf16", k_near: float = 2.0, pattern=None,
                 split: bool = False, cast_chunks: int = 1):
        if storage not in ("bf16", "fp32", "int8"):
            raise ValueError(f"storage {storage!r} (expected 'bf16', 'fp32' or 'int8')")
        if split and storage != "int8":
            raise ValueError("the split far field (codes + remainders) is an int8 storage")
        _lib.require_gpu()
        A = assembled.A_dev
        n = A.shape[0]
        dev = A.device
        self.n = n
        self.bf16 = storage == "bf16"
        self.i8 = storage == "int8"
        self.storage = storage
        t0 = time.perf_counter()
        rowptr, col = pattern if pattern is not None else near_structure(assembled, k_near)
        self.nbr = rowptr.size - 1
        self.nnzb = int(rowptr[-1])
        self.rowptr_h, self.col_h = rowptr, col
        # pinned, non-blocking uploads: nothing here waits for the device
        self.rowptr = _lib.upload(rowptr, I64)
        self.col = _lib.upload(col, I64)
        self.vals = torch.empty(9 * max(self.nnzb, 1), dtype=F64, device=dev)
        align = {"bf16": 8, "fp32": 4, "int8": 16}[storage]
        self.ldf = ((n + align - 1) // align) * align
        if self.i8:
            # near blocks held out of A (saved to vals, zeroed) so the segment
            # scales see the far field only, then written back
            self.nseg = int(_lib.load().mbem_i8_segments(n))
            self.F = torch.empty((n, self.ldf), dtype=torch.uint8, device=dev)
            self.S = torch.empty((n, max(self.nseg, 1)), dtype=torch.float32, device=dev)
            # with split, each code's rounding remainder too (24 bits, packed
            # four to three words): the outer operator of SplitOuter
            self.lde = int(_lib.load().mbem_i8e_resid_bytes(n))
            self.E = torch.empty((n, self.lde), dtype=torch.uint8, device=dev) if split else None
            hold = (self.nbr, ptr(self.rowptr), ptr(self.col), self.nnzb, ptr(self.vals))
            call("mbem_near_hold", ptr(A), A.stride(0), *hold, 0, stream())
            # the cast in row chunks, an event after each (``cast_events``):
            # a consumer of the codes (the coarse image) can follow chunk by
            # chunk on another stream while the cast goes on
            self.cast_events = []
            bounds = np.linspace(0, n, max(1, int(cast_chunks)) + 1).astype(np.int64)
            try:
                for r0, r1 in zip(bounds[:-1], bounds[1:]):
                    r0, r1 = int(r0), int(r1)
                    if r1 <= r0:
                        continue
                    if split:
                        call("mbem_cast_i8e", ptr(A[r0:r1]), A.stride(0), r1 - r0, n,
                             ptr(self.F[r0:r1]), self.ldf, ptr(self.S[r0:r1]), self.nseg,
                             ptr(self.E[r0:r1]), self.lde, stream())
                    else:
                        call("mbem_cast_i8", ptr(A[r0:r1]), A.stride(0), r1 - r0, n,
                             ptr(self.F[r0:r1]), self.ldf, ptr(self.S[r0:r1]), self.nseg,
                             stream())
                    ev = torch.cuda.Event()
                    ev.record()
                    self.cast_events.append((r0, r1, ev))
            finally:
                # the near blocks go back into the caller's A whatever the
                # cast returned: A is never left with them zeroed
                call("mbem_near_hold", ptr(A), A.stride(0), *hold, 1, stream())
            self.far_bytes = n * n + n * self.nseg * 4
        else:
            dt = torch.int16 if self.bf16 else torch.float32
            self.F = torch.empty((n, self.ldf), dtype=dt, device=dev)
            call("mbem_cast_lo", ptr(A), A.stride(0), n, n, ptr(self.F), self.ldf, int(self.bf16),
                 stream())
            call("mbem_near_extract", ptr(A), A.stride(0), self.nbr, ptr(self.rowptr),
                 ptr(self.col), self.nnzb, ptr(self.vals), ptr(self.F), self.ldf, int(self.bf16),
                 stream())
            self.far_bytes = n * n * (2 if self.bf16 else 4)
            self.E = None
            self.cast_events = []
        self._lower = {}
        self.xf = torch.empty(max(n, 4), dtype=torch.float32, device=dev)   # FP32 copy of x
        # host time only: the casts are queued, not waited for (stream order
        # puts them ahead of any use)
        self.build_s = time.perf_counter() - t0
        self.near_bytes = self.nnzb * (72 + 8)

    def _far(self, r0, r1, ncols, alpha, beta, out):
        """out[:] = alpha F[r0:r1, :ncols] xf[:ncols] + beta out (xf already set)."""
        if self.i8:
            call("mbem_gemv_i8", ptr(self.F[r0:r1]), self.ldf, ptr(self.S[r0:r1]), self.nseg,
                 r1 - r0, ncols, ptr(self.xf), ptr(out), alpha, beta, stream())
        else:
            call("mbem_gemv_lo", ptr(self.F[r0:r1]), self.ldf, r1 - r0, ncols, ptr(self.xf),
                 ptr(out), alpha, beta, int(self.bf16), stream())

    def apply_into(self, v, out):
        call("mbem_vec_to_f32", ptr(v), self.n, ptr(self.xf), stream())
        self._far(0, self.n, self.n, 1.0, 0.0, out)
        call("mbem_bsr_gemv", self.nbr, ptr(self.rowptr), ptr(self.col), ptr(self.vals), ptr(v),
             ptr(out), 1.0, stream())
        return out

    def matvec(self, v):
        out = torch.empty_like(v)
        return self.apply_into(v, out)

    def _lower_bsr(self, r0, r1):
        """Near blocks of rows [r0, r1) with columns < r0 (block-CSR)."""
        key = (r0, r1)
        if key not in self._lower:
            b0, b1, c_lim = r0 // 3, r1 // 3, r0 // 3
            rp = self.rowptr_h[b0:b1 + 1]
            keep = self.col_h[rp[0]:rp[-1]] < c_lim
            idx = np.arange(rp[0], rp[-1])[keep]
            row_of = np.repeat(np.arange(b1 - b0), np.diff(rp))
            counts = np.bincount(row_of[keep], minlength=b1 - b0).astype(np.int64)
            rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
            dev = self.vals.device
            gather = torch.as_tensor(idx, dtype=I64, device=dev)
            vals = self.vals.view(-1, 9)[gather].reshape(-1).contiguous() if idx.size else \
                torch.zeros(9, dtype=F64, device=dev)
            self._lower[key] = (b1 - b0, torch.as_tensor(rowptr, dtype=I64, device=dev),
                                torch.as_tensor(self.col_h[idx], dtype=I64, device=dev), vals)
        return self._lower[key]

    def lower_into(self, r0, r1, z, out):
        """out -= (F + N)[r0:r1, :r0] z[:r0]: the Gauss-Seidel lower coupling."""
        call("mbem_vec_to_f32", ptr(z), r0, ptr(self.xf), stream())
        self._far(r0, r1, r0, -1.0, 1.0, out)
        nbr, rowptr, col, vals = self._lower_bsr(r0, r1)
        call("mbem_bsr_gemv", nbr, ptr(rowptr), ptr(col), ptr(vals), ptr(z), ptr(out), -1.0,
             stream())


class SplitOuter:
    """The outer loop's operator from the split far field: A v as
    s (u - 128 + e 2^-24) v over the far entries (int8 codes u and 24-bit
    remainders e, FP64 v and FP64 sums) plus the FP64 near blocks.  Every far
    entry is A's to within s 2^-25 (s = its segment's largest magnitude /
    127, so about 2e-10 of that magnitude; the near blocks are exact), read
    from 4 bytes instead of 8: the FP64 pass over A in each outer step of the
    solve (la/solver.py:178) becomes a pass over 5/8 of the bytes."""

    def __init__(self, op: MixedOperator):
        if op.E is None:
            raise ValueError("SplitOuter needs a MixedOperator built with split=True")
        self.op = op
        self.n = op.n
        self.bytes = op.n * op.ldf + op.n * op.lde + op.n * op.nseg * 4 + op.near_bytes

    def apply_into(self, v, out):
        o = self.op
        call("mbem_gemv_i8e", ptr(o.F), o.ldf, ptr(o.E), o.lde, ptr(o.S), o.nseg, o.n, o.n,
             ptr(v), ptr(out), 1.0, 0.0, stream())
        call("mbem_bsr_gemv", o.nbr, ptr(o.rowptr), ptr(o.col), ptr(o.vals), ptr(v), ptr(out),
             1.0, stream())
        return out

    def matvec(self, v):
        return self.apply_into(v, torch.empty_like(v))


class _InnerSolve:
    """Right preconditioner of the outer loop: an inner FGMRES on the mixed
    operator to a loose tolerance."""

    def __init__(self, op, M, rtol, maxiter):
        self.op, self.M, self.rtol, self.maxiter = op, M, rtol, maxiter
        self.iterations = []

    def apply_into(self, r, z):
        x, rep = fgmres(self.op, r, M=self.M, rtol=self.rtol, restart=self.maxiter,
                        maxiter=self.maxiter)
        self.iterations.append(rep.iterations)
        z.copy_(x)
        return z


class _RecycledInner:
    """Inner solve with a recycled subspace (GCRO with a fixed recycle space;
    the reference's GCRO-DR is solver.py:248-395).  The inner operator is the
    same at every outer step, so each inner solve keeps its Krylov space as
    pairs (U, C) with op U = C and C^T C = I; the next solve first takes the
    part of its residual in span C (x0 = U C^T r) and then runs flexible
    GMRES on (I - C C^T) op, minimising over span U + the new Krylov space.
    After the solve, op (Z - U Hc) = V Hbar (Hbar unrotated) gives the new
    pairs: Hbar = Q R, C_new = V Q, U_new = (Z - U Hc) R^-1, appended until
    ``kmax`` vectors are held."""

    def __init__(self, op, M, rtol, maxiter, n, dev, kmax=60):
        self.op, self.M = op, M
        self.rtol, self.maxiter = rtol, maxiter
        self.kmax = int(kmax)
        self.U = torch.zeros((self.kmax, n), dtype=F64, device=dev)
        self.C = torch.zeros((self.kmax, n), dtype=F64, device=dev)
        self.k = 0
        # the C-projection coefficients of an Arnoldi step sit in the tail of
        # the CGS2 buffer and reach the host with the step's own coefficients
        self.ws = _Krylov(n, maxiter, dev, tail=max(self.kmax, 1))
        self.cwork = torch.empty(int(_lib.load().mbem_multidot_work(max(self.kmax, 1), n)),
                                 dtype=F64, device=dev)
        self.iterations = []
        self.n = n

    def apply_into(self, r, z):
        ws, n, k = self.ws, self.n, self.k
        m = self.maxiter
        x = z
        x.zero_()
        rr = r.clone()
        # ||r|| and, after the projection onto span C, ||rr|| reach the host
        # in one transfer (the projection of a zero r is zero)
        ws.dots(rr.view(1, -1), 1, rr, 0)
        if k > 0:
            coef = torch.empty(k, dtype=F64, device=rr.device)
            call("mbem_multidot", ptr(self.C), self.C.stride(0), k, n, ptr(rr), ptr(coef),
                 ptr(self.cwork), stream())
            ws.axpy_rows(self.U, k, coef, x, 1.0)
            ws.axpy_rows(self.C, k, coef, rr, -1.0)
        ws.dots(rr.view(1, -1), 1, rr, 1)
        hv = ws.h[:2].cpu().numpy()
        r_norm = float(np.sqrt(max(float(hv[0]), 0.0)))
        if r_norm == 0.0:
            self.iterations.append(0)
            return z
        beta = float(np.sqrt(max(float(hv[1]), 0.0)))
        V, Z = ws.V, ws.Z
        H = np.zeros((m + 1, m))
        Hraw = np.zeros((m + 1, m))
        cs, sn, g = np.zeros(m), np.zeros(m), np.zeros(m + 1)
        Hc_rows = np.zeros((m, k))
        j_done = 0
        if beta / r_norm > self.rtol:
            axpby(1.0 / beta, rr, 0.0, V[0])
            g[0] = beta
            for j in range(m):
                if j == 0:
                    self.M.apply_into(V[j], Z[j])
                w = V[j + 1]
                self.op.apply_into(Z[j], w)
                if k > 0:
                    hc = ws.tail_view(k)
                    call("mbem_multidot", ptr(self.C), self.C.stride(0), k, n, ptr(w), ptr(hc),
                         ptr(self.cwork), stream())
                    ws.axpy_rows(self.C, k, hc, w, -1.0)
                # the next step's preconditioner apply is queued before the
                # host waits for this step's coefficients (it only needs the
                # normalized V[j + 1], already queued); unused if the solve
                # stops here.  Z has m rows: the last step queues none.
                nxt = (lambda j=j: self.M.apply_into(V[j + 1], Z[j + 1])) if j + 1 < m else None
                if k > 0:
                    hsum, wnorm, Hc_rows[j] = ws.cgs2(V, j + 1, w, normalize=True, tail=k,
                                                      before_wait=nxt)
                else:
                    hsum, wnorm = ws.cgs2(V, j + 1, w, normalize=True, before_wait=nxt)
                H[:j + 1, j] = hsum
                H[j + 1, j] = wnorm
                Hraw[:j + 2, j] = H[:j + 2, j]
                for i in range(j):
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                d = np.hypot(H[j, j], H[j + 1, j])
                j_done = j + 1
                if d == 0.0:
                    break
                cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
                H[j, j] = d
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                if abs(g[j + 1]) / r_norm < self.rtol or wnorm <= 1e-300:
                    break
        self.iterations.append(j_done)
        if j_done == 0:
            return z
        y = _solve_upper(H[:j_done, :j_done], g[:j_done])
        Hc = Hc_rows[:j_done].T if k > 0 else None                        # k x j_done
        # x += Z y - U (Hc y): one upload of both coefficient vectors
        coefs = np.concatenate([y, Hc @ y]) if k > 0 else y
        cd = _lib.upload(coefs, F64)                        # pinned, no host wait
        ws.axpy_rows(Z, j_done, cd[:j_done], x, 1.0)
        if k > 0:
            ws.axpy_rows(self.U, k, cd[j_done:], x, -1.0)
        self._grow(j_done, Hraw[:j_done + 1, :j_done], Hc)
        return z

    def _grow(self, j, Hbar, Hc):
        """Append the pairs of this solve's Krylov space to (U, C)."""
        add = min(j, self.kmax - self.k)
        if add <= 0:
            return
        Q, R = np.linalg.qr(Hbar)                  # (j+1) x j, j x j
        if np.min(np.abs(np.diag(R))) < 1e-12 * np.max(np.abs(np.diag(R))):
            return
        Rinv = np.linalg.inv(R)
        ws, k, dev = self.ws, self.k, self.U.device
        V, Z = ws.V, ws.Z
        # one upload of every coefficient column: rows of Q^T, Rinv^T, Hc^T
        parts = [Q[:, :add].T, Rinv[:, :add].T] + ([Hc.T] if Hc is not None and k > 0 else [])
        cf = _lib.upload(np.concatenate([p.reshape(-1) for p in parts]), F64)
        q_t = cf[:add * (j + 1)].view(add, j + 1)
        r_t = cf[add * (j + 1):add * (2 * j + 1)].view(add, j)
        n = self.n
        # three row-major FP64 products instead of a launch per vector (the
        # per-vector form was ~5 j host-bound launches per solve)
        if Hc is not None and k > 0:
            h_t = cf[add * (2 * j + 1):].view(j, k)
            W = torch.empty((j, n), dtype=F64, device=dev)
            W.copy_(Z[:j])
            call("mbem_dgemm_rowmajor", j, n, k, -1.0, ptr(h_t), k, ptr(self.U),    # W = Z - Hc^T U
                 self.U.stride(0), 1.0, ptr(W), W.stride(0), stream())
        else:
            W = Z[:j]
        call("mbem_dgemm_rowmajor", add, n, j + 1, 1.0, ptr(q_t), j + 1, ptr(V), V.stride(0),
             0.0, ptr(self.C[k]), self.C.stride(0), stream())           # C_new = Q^T V
        call("mbem_dgemm_rowmajor", add, n, j, 1.0, ptr(r_t), j, ptr(W), W.stride(0),
             0.0, ptr(self.U[k]), self.U.stride(0), stream())           # U_new = R^-T W
        self.k = k + add


@dataclass
class MixedReport:
    outer: SolveReport
    inner_iterations: list = field(default_factory=list)
    operator_build_s: float = 0.0
    precond_build_s: float = 0.0
    solve_s: float = 0.0
    storage: str = "bf16"
    near_blocks: int = 0
    far_bytes: int = 0
    coarse: dict | None = None
    outer_operator: str = "fp64"


    @property
    def converged(self):
        return self.outer.converged

    @property
    def iterations(self):
        return self.outer.iterations

    def __str__(self):
        c = "" if not self.coarse else (f", coarse {self.coarse['basis']} rank "
                                        f"{self.coarse['rank']}")
        o = ("FP64 outer" if self.outer_operator == "fp64" else
             "outer on int8 codes + 24-bit remainders, FP64 sums")
        return (f"mixed ({self.storage} far field, FP64 near, {o}{c}): outer {self.outer}, "
                f"inner iterations {self.inner_iterations}")


_SIDE = {}
# row chunks of the int8 cast when a coarse space follows it (its image of
# each chunk runs beside the cast of the next)
CAST_CHUNKS = 6


def _side_stream(k: int = 0):
    """Side stream ``k`` of the operator build, one per device for the life
    of the process (default priority): 0 for the operator, 1 for the coarse
    space.  A new stream per solve would leave the caching allocator's blocks
    of the last build (held per stream) unusable by the next, so every solve
    would map its far-field copy afresh."""
    key = (torch.cuda.current_device(), k)
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(device=key[0])
    return _SIDE[key]


def solve_mixed(assembled, rtol: float = 1e-10, storage: str = "bf16", k_near: float = 2.0,
                inner_rtol: float = 5e-2, inner_maxiter: int = 60, outer_restart: int = 30,
                outer_maxiter: int = 40, precond_max_dense=0, precond_bj_chunk=1500,
                precond_grouping: str = "region", pattern=None, recycle: int = 90,
                coarse: str = None, outer: str = "fp64"):
    """Mixed-precision FGMRES of ``assembled``; returns (x device, MixedReport).
    ``recycle`` > 0 keeps up to that many recycled vectors across the inner
    solves (``_RecycledInner``); ``coarse`` ("const", "rigid", "linear")
    adds the coarse correction of la/coarse.py to the inner block-Jacobi.
    ``outer`` is the outer loop's operator: "fp64" the assembled FP64 matrix,
    "split" (int8 storage) ``SplitOuter``, the int8 codes plus their 24-bit
    rounding remainders (every far entry within 2e-10 of its segment's
    largest, FP64 sums, near blocks exact), half the FP64 bytes per outer
    step."""
    if outer not in ("fp64", "split"):
        raise ValueError(f"outer {outer!r} (expected 'fp64' or 'split')")
    if outer == "split" and storage != "int8":
        raise ValueError("outer='split' needs storage='int8'")
    from .preconditioner import BlockGaussSeidel
    # Three streams, all at the default priority.  The near pattern needs
    # the geometry only: its counting pass is queued on a side stream beside
    # the assembly and read back after the block-Jacobi host work, which
    # runs while the assembly occupies the device.  Once the block-Jacobi
    # gathers have read A (``gathered``) and the chunk inversions are queued
    # on the main stream, the operator build goes on the side stream (near
    # blocks held out of A, the int8 cast in row chunks, blocks restored),
    # and the coarse space on a second side stream, its image following the
    # cast chunk by chunk; the inversions work on their own copies of the
    # chunks.
    t0 = time.perf_counter()
    main = torch.cuda.current_stream()
    side = _side_stream()
    near = None
    if pattern is None:
        with torch.cuda.stream(side):
            near = NearPattern(assembled, k_near)     # counting pass queued, not waited for
    built = {}

    i8 = storage == "int8"
    side2 = _side_stream(1) if coarse else None

    def start_operator(gathered, chunk_dofs):
        pat = near.result() if near is not None else pattern
        side.wait_event(gathered)
        with torch.cuda.stream(side):
            built["op"] = MixedOperator(assembled, storage=storage, k_near=k_near, pattern=pat,
                                        split=outer == "split",
                                        cast_chunks=CAST_CHUNKS if (coarse and i8) else 1)
        if coarse:
            # the coarse space on a second side stream: the image of the
            # mixed operator itself (int8 codes and near blocks) follows the
            # cast chunk by chunk, so most of it runs under the HBM-bound
            # cast; for BF16 / FP32 storage it is the FP64 A's image, which
            # waits for the build to put A's near blocks back
            from .coarse import CoarseSpace
            if not i8:
                side2.wait_stream(side)
            with torch.cuda.stream(side2):
                built["cs"] = CoarseSpace(assembled, chunk_dofs, basis=coarse,
                                          op=built["op"]).finish()

    M = BlockGaussSeidel(assembled, max_dense=precond_max_dense, bj_chunk=precond_bj_chunk,
                         grouping=precond_grouping, inverse_dtype="f32", defer_check=True,
                         after_gather=start_operator)
    op = built["op"]
    cs = built.get("cs")
    t1 = time.perf_counter()
    main.wait_stream(side)
    if side2 is not None:
        main.wait_stream(side2)
    M.coupling = op                      # the lower couplings read the mixed operator
    M.finish()                           # one wait for the assembly and both builds
    P = M
    if cs is not None:
        from .coarse import TwoLevel
        P = TwoLevel(M, cs.finish())
    t2 = time.perf_counter()
    if recycle:
        inner = _RecycledInner(op, P, inner_rtol, inner_maxiter, op.n, assembled.A_dev.device,
                               kmax=recycle)
    else:
        inner = _InnerSolve(op, P, inner_rtol, inner_maxiter)
    A_outer = SplitOuter(op) if outer == "split" else assembled
    x, outer_rep = fgmres(A_outer, assembled.b_dev, M=inner, rtol=rtol, restart=outer_restart,
                          maxiter=outer_maxiter)
    torch.cuda.current_stream().synchronize()
    t3 = time.perf_counter()
    rep = MixedReport(outer=outer_rep, inner_iterations=inner.iterations, operator_build_s=t2 - t1,
                      precond_build_s=t1 - t0, solve_s=t3 - t2, storage=storage,
                      near_blocks=op.nnzb, far_bytes=op.far_bytes,
                      coarse=(P.summary()["coarse"] if coarse else None), outer_operator=outer)
    return x, rep
