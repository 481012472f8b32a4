"""K independent FGMRES runs in lockstep on one pair of geometry operators.

System c is B(g_c) y_c = b_c with B(g) = A0 + g A1 (the Laplace-state form of
a relaxing region, see ``mbem_amd.viscoelastic``).  Every block vector is a
row-major n x K tensor whose column c belongs to run c, so one matrix-core
GEMM (``mbem_gemm_g``) applies all K operators at once: A0 and A1 are read
once per iteration for all runs.

The algorithm of each run is the reference's right-preconditioned flexible
GMRES (solver.py:115-245): Arnoldi with classical Gram-Schmidt applied twice
(CGS2, in place of modified Gram-Schmidt: the same subspace in exact
arithmetic, and one pass over the basis per sweep), Givens rotations,
x += Z y at a restart, and the true residual to decide convergence.  The
Givens updates and the triangular solve run on the device, one thread per
run, so a restart cycle is a fixed kernel sequence; with ``graph=True`` it
is captured once into HIP graphs (one per segment of ``segment`` Arnoldi
steps, plus the end of cycle) and replayed.  The host reads the K residual
estimates after each segment and stops the cycle early once every run has
met the tolerance.

Preconditioner: block-Jacobi over the cluster chunks of the (permuted)
operator, one set of FP64 chunk inverses per anchor value of g; each run
uses the inverses of the anchor nearest its g (FGMRES allows any fixed
preconditioner per run).  Applied as one grouped GEMM over a work table.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib
from .._lib import call, ptr, stream
from . import autotune

F64 = torch.float64
I64 = torch.int64


class GOperator:
    """Y = alpha B(g) X + beta Y for n x K blocks, B(g_c) = A0 + g_c A1.

    Candidates: the compiled tiles of the project's matrix-core kernel in
    its two-operator form (``mbem_gemm_g``, ids 0..C-1) and, when A0 and A1
    sit side by side in one n x 2L row-major buffer (``pair=(AA, L)``:
    A0 = AA[:, :n], A1 = AA[:, L:L+n]), the one GEMM [A0 | A1] [X ; X diag(g)]
    through the library DGEMM (``LIB`` = C) or through the project's kernel
    in its one-operator form (``mbem_gemm_rm``, ids C+1+cfg).  ``tune`` times
    them on the real shape and keeps the fastest."""

    def __init__(self, A0: torch.Tensor, A1: torch.Tensor, g: torch.Tensor, pair=None):
        n = A0.shape[0]
        if A0.shape != (n, n) or A1.shape != (n, n) or A0.stride(0) != A1.stride(0):
            raise ValueError("A0 and A1 must be n x n with one leading dimension")
        if A0.stride(0) % 2:
            raise ValueError("the leading dimension of A0/A1 must be even (16-byte rows)")
        self.A0, self.A1, self.n = A0, A1, n
        self.g = g.contiguous()
        self.K = self.g.numel()
        self.cfg = 0
        self.LIB = len(autotune.configs())
        self.pair = pair
        self.X2 = None
        if pair is not None:
            AA, L = pair
            if AA.shape != (n, 2 * L) or AA.data_ptr() != A0.data_ptr() or \
                    A1.data_ptr() != AA[:, L:].data_ptr():
                raise ValueError("pair must be the n x 2L buffer holding A0 and A1")

    def candidates(self):
        lib = _lib.load()
        cfgs = [i for i, _ in autotune.configs()]
        c = [i for i in cfgs if lib.mbem_mgemm_config_fits(i, 1)]
        if self.pair is None:
            return c
        return c + [self.LIB] + [self.LIB + 1 + i for i in cfgs if lib.mbem_mgemm_config_fits(i, 0)]

    def describe(self, cfg=None) -> str:
        """What candidate ``cfg`` (default: the chosen one) is."""
        cfg = self.cfg if cfg is None else cfg
        tiles = dict(autotune.configs())
        if cfg == self.LIB:
            return "library DGEMM on [A0 | A1] [X ; X g]"
        t = tiles[cfg - self.LIB - 1] if cfg > self.LIB else tiles[cfg]
        form = "one GEMM on [A0 | A1] [X ; X g]" if cfg > self.LIB else "two-operator form"
        return (f"project kernel {t['bm']}x{t['bn']}x{t['bk']} tile, {t['wm']}x{t['wn']} waves, "
                f"{form}")

    def tune(self, X, Y):
        """Pick the configuration for this shape (times every candidate)."""
        key = f"gemm_g n={self.n} K={self.K} lib={self.pair is not None}"
        self._prepare()
        self.cfg = autotune.tune(key, lambda cfg: self._launch(X, Y, 1.0, 0.0, cfg),
                                 candidates=self.candidates())
        return self.cfg

    def _prepare(self):
        """The library path's [X ; X g] buffer (made before any capture)."""
        if self.pair is not None and self.X2 is None:
            L = self.pair[1]
            self.X2 = torch.zeros((2 * L, self.K), dtype=F64, device=self.A0.device)

    def _launch(self, X, Y, alpha, beta, cfg):
        if cfg >= self.LIB:
            AA, L = self.pair
            n, K = self.n, self.K
            call("mbem_bcopy", ptr(X), ptr(self.X2), n * K, stream())
            call("mbem_bscale", ptr(X), ptr(self.g), ptr(self.X2[L:]), n, K, stream())
            if cfg == self.LIB:
                call("mbem_dgemm_rowmajor", n, K, 2 * L, float(alpha), ptr(AA), AA.stride(0),
                     ptr(self.X2), K, float(beta), ptr(Y), Y.stride(0), stream())
            else:
                call("mbem_gemm_rm", ptr(AA), AA.stride(0), n, 2 * L, ptr(self.X2), K, K, ptr(Y),
                     Y.stride(0), float(alpha), float(beta), int(cfg - self.LIB - 1), stream())
            return
        call("mbem_gemm_g", ptr(self.A0), ptr(self.A1), self.A0.stride(0), self.n, self.n, ptr(X),
             X.stride(0), self.K, ptr(self.g), ptr(Y), Y.stride(0), float(alpha), float(beta),
             int(cfg), stream())

    def apply(self, X, Y, alpha=1.0, beta=0.0):
        self._launch(X, Y, alpha, beta, self.cfg)
        return Y

    @property
    def flops(self) -> float:
        """Algorithmic flops of one apply: two n x n x K GEMMs."""
        return 4.0 * self.n * self.n * self.K


class GroupedBlockJacobi:
    """Z = M(R) on n x K blocks: chunk k of the rows (contiguous, cluster
    order) times the inverse of B(g_a)'s diagonal chunk, for the column range
    of anchor a."""

    def __init__(self, A0, A1, chunk_sizes, anchors, col_ranges, K):
        dev = A0.device
        n = A0.shape[0]
        self.n, self.K = n, K
        self.sizes = [int(m) for m in chunk_sizes]
        if sum(self.sizes) != n:
            raise ValueError("chunk sizes must cover the operator")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])[:-1].astype(np.int64)
        self.anchors = [float(a) for a in anchors]
        self.col_ranges = [(int(a), int(b)) for a, b in col_ranges]
        for c0, c1 in self.col_ranges:
            if c0 % 2:
                raise ValueError("anchor column ranges must start at even columns")
        t0 = time.perf_counter()
        self.ld = [m + (m & 1) for m in self.sizes]
        per_anchor = sum(m * ld for m, ld in zip(self.sizes, self.ld))
        self.inv = torch.zeros(len(self.anchors) * per_anchor + 2, dtype=F64, device=dev)
        self.a_off = np.zeros((len(self.anchors), len(self.sizes)), dtype=np.int64)
        idx_all = torch.arange(n, dtype=I64, device=dev)
        offs = np.concatenate([[0], np.cumsum([m * ld for m, ld in zip(self.sizes, self.ld)])])
        sizes = np.asarray(self.sizes)
        order = np.argsort(sizes, kind="stable")
        for a, g in enumerate(self.anchors):
            base = a * int(offs[-1])
            self.a_off[a] = base + offs[:-1]
            # size classes (largest / smallest <= 1.1), each padded with an
            # identity block to its largest size and inverted in one batch
            c0 = 0
            while c0 < len(order):
                c1 = c0
                while c1 + 1 < len(order) and sizes[order[c1 + 1]] <= 1.1 * sizes[order[c0]]:
                    c1 += 1
                cls = order[c0:c1 + 1]
                mp, cnt = int(sizes[cls].max()), len(cls)
                tmp = torch.empty((cnt, mp, mp), dtype=F64, device=dev)
                for q, k in enumerate(cls):
                    o, m = int(self.offsets[k]), int(sizes[k])
                    call("mbem_gather_comb_pad", ptr(A0), ptr(A1), A0.stride(0),
                         ptr(idx_all[o:o + m]), m, mp, g, ptr(tmp[q]), stream())
                ipiv = torch.empty((cnt, mp), dtype=torch.int32, device=dev)
                info = torch.empty(cnt, dtype=torch.int32, device=dev)
                call("mbem_invert_batched", ptr(tmp), mp, cnt, ptr(ipiv), ptr(info), stream())
                if int(info.abs().max().item()) != 0:
                    raise np.linalg.LinAlgError(f"block-Jacobi chunk of B(g={g:.6g}) is singular")
                for q, k in enumerate(cls):
                    m = int(sizes[k])
                    call("mbem_copy2d", ptr(tmp[q]), mp, ptr(self.inv[int(self.a_off[a, k]):]),
                         self.ld[k], m, m, 8, stream())
                del tmp
                c0 = c1 + 1
        off = len(self.anchors) * int(offs[-1])
        torch.cuda.current_stream().synchronize()
        self.build_s = time.perf_counter() - t0
        self.inverse_bytes = 8 * off
        self._tables = {}
        self.cfg = 0

    def _table(self, cfg):
        if cfg not in self._tables:
            c = dict(autotune.configs())[cfg]
            bm, bn = c["bm"], c["bn"]
            rows = []
            for a, (c0, c1) in enumerate(self.col_ranges):
                if c1 <= c0:
                    continue
                tc = (c1 - c0 + bn - 1) // bn
                for k, m in enumerate(self.sizes):
                    tr = (m + bm - 1) // bm
                    base = int(self.offsets[k]) * self.K + c0
                    for i in range(tr):
                        for j in range(tc):
                            rows.append((self.a_off[a, k], self.ld[k], m, base, c1 - c0, i, j, 0))
            w = np.asarray(rows, dtype=np.int64).reshape(-1, 8)
            if _lib.load().mbem_work_item_bytes() != 64:
                raise RuntimeError("work item layout mismatch")
            self._tables[cfg] = (torch.as_tensor(w, device=self.inv.device), w.shape[0])
        return self._tables[cfg]

    def tune(self, R, Z):
        key = f"grouped_bj n={self.n} K={self.K} chunks={len(self.sizes)} anchors={len(self.anchors)}"
        self.cfg = autotune.tune(key, lambda cfg: self._launch(R, Z, cfg))
        return self.cfg

    def _launch(self, R, Z, cfg):
        tab, nw = self._table(cfg)
        call("mbem_gemm_grouped", ptr(self.inv), ptr(R), R.stride(0), ptr(Z), Z.stride(0), ptr(tab),
             nw, int(cfg), stream())

    def apply(self, R, Z):
        self._launch(R, Z, self.cfg)
        return Z


@dataclass
class BatchedReport:
    converged: np.ndarray
    iterations: np.ndarray          # Arnoldi steps until each run met the tolerance
    cycles: int
    relres: np.ndarray              # true relative residuals at the end
    seconds: float
    graph: bool
    steps: int                      # Arnoldi steps executed (all runs in lockstep)
    timings: dict = field(default_factory=dict)

    def __str__(self):
        return (f"batched FGMRES: {int(self.converged.sum())}/{self.converged.size} converged, "
                f"{self.steps} lockstep steps in {self.cycles} cycle(s), max relres "
                f"{float(self.relres.max()):.3e}, {self.seconds:.3f} s"
                f"{' (HIP graphs)' if self.graph else ''}")


class BatchedFGMRES:
    """Lockstep FGMRES(m) of K runs: ``solve(B)`` for an n x K right-hand side."""

    NSLOTS = 256          # row slices of the block multi-dot (a multiple of 4)

    def __init__(self, op: GOperator, M: GroupedBlockJacobi, restart: int = 30,
                 segment: int = 5, graph: bool = True):
        self.op, self.M = op, M
        self.n, self.K = op.n, op.K
        self.m = int(restart)
        self.segment = max(1, int(segment))
        self.graph = graph
        dev = op.A0.device
        n, K, m = self.n, self.K, self.m
        if K % 2:
            raise ValueError("K must be even (16-byte block rows); pad with a zero column")
        e = dict(dtype=F64, device=dev)
        self.V = torch.zeros((m + 1, n, K), **e)
        self.Z = torch.zeros((m, n, K), **e)
        self.X = torch.zeros((n, K), **e)
        self.R = torch.zeros((n, K), **e)
        self.B = torch.zeros((n, K), **e)
        self.H = torch.zeros((K, m + 1, m), **e)
        self.cs = torch.zeros((K, m), **e)
        self.sn = torch.zeros((K, m), **e)
        self.gvec = torch.zeros((K, m + 1), **e)
        self.y = torch.zeros((K, m), **e)
        self.hp1 = torch.zeros((K, m + 1), **e)
        self.hp2 = torch.zeros((K, m + 1), **e)
        self.sumsq = torch.zeros(K, **e)
        self.beta = torch.zeros(K, **e)
        self.scale = torch.zeros(K, **e)
        self.res = torch.zeros(K, **e)
        self.partial = torch.zeros(self.NSLOTS * 8 * K, **e)
        self.jdone = torch.zeros(1, dtype=torch.int32, device=dev)
        self.res_host = torch.zeros(K, dtype=F64, pin_memory=True)
        self.beta_host = torch.zeros(K, dtype=F64, pin_memory=True)
        self._graphs = {}
        self.bstride = n * K

    # -- kernel sequences ------------------------------------------------------
    def _norms(self, W):
        """sumsq[c] = |W[:, c]|^2"""
        call("mbem_bdot", ptr(W), self.bstride, 1, ptr(W), self.n, self.K, ptr(self.partial),
             self.NSLOTS, stream())
        call("mbem_breduce", ptr(self.partial), self.NSLOTS, 1, self.K, ptr(self.sumsq), 1, 0, 0,
             None, stream())

    def _project(self, j, hp):
        """hp[c, 0..j] = V_i^T w (columnwise); w -= sum_i hp[c, i] V_i."""
        w = self.V[j + 1]
        for i0 in range(0, j + 1, 8):
            cnt = min(8, j + 1 - i0)
            call("mbem_bdot", ptr(self.V[i0]), self.bstride, cnt, ptr(w), self.n, self.K,
                 ptr(self.partial), self.NSLOTS, stream())
            call("mbem_breduce", ptr(self.partial), self.NSLOTS, cnt, self.K, ptr(hp), self.m + 1,
                 i0, 0, None, stream())
        call("mbem_baxpy", ptr(self.V[0]), self.bstride, j + 1, None, ptr(hp), self.m + 1, -1.0,
             ptr(w), self.n, self.K, stream())

    def _step(self, j):
        self.M.apply(self.V[j], self.Z[j])
        self.op.apply(self.Z[j], self.V[j + 1])
        self._project(j, self.hp1)
        self._project(j, self.hp2)
        self._norms(self.V[j + 1])
        call("mbem_fg_givens", self.K, self.m, j, ptr(self.hp1), ptr(self.hp2), ptr(self.H),
             ptr(self.cs), ptr(self.sn), ptr(self.gvec), ptr(self.sumsq), ptr(self.scale),
             ptr(self.res), ptr(self.jdone), stream())
        call("mbem_bscale", ptr(self.V[j + 1]), ptr(self.scale), ptr(self.V[j + 1]), self.n,
             self.K, stream())

    def _start(self):
        """V0 = R / |R|, gvec = |R| e1, jdone = 0 (R holds the residual)."""
        self._norms(self.R)
        call("mbem_fg_start", self.K, self.m, ptr(self.sumsq), ptr(self.beta), ptr(self.scale),
             ptr(self.gvec), ptr(self.jdone), stream())
        call("mbem_bscale", ptr(self.R), ptr(self.scale), ptr(self.V[0]), self.n, self.K, stream())
        call("mbem_memcpy_async", ptr(self.beta_host), ptr(self.beta), 8 * self.K, stream())

    def _end(self):
        """X += Z y; R = B - B(g) X; then the next cycle's start."""
        call("mbem_fg_solve", self.K, self.m, ptr(self.H), ptr(self.gvec), ptr(self.jdone),
             ptr(self.y), stream())
        call("mbem_baxpy", ptr(self.Z[0]), self.bstride, self.m, ptr(self.jdone), ptr(self.y),
             self.m, 1.0, ptr(self.X), self.n, self.K, stream())
        call("mbem_bcopy", ptr(self.B), ptr(self.R), self.n * self.K, stream())
        self.op.apply(self.X, self.R, alpha=-1.0, beta=1.0)
        self._start()

    def _segment(self, q):
        j0 = q * self.segment
        for j in range(j0, min(j0 + self.segment, self.m)):
            self._step(j)
        call("mbem_memcpy_async", ptr(self.res_host), ptr(self.res), 8 * self.K, stream())

    def _run(self, key, fn):
        if not self.graph:
            fn()
            return
        if key not in self._graphs:
            call("mbem_graph_begin", stream())
            try:
                fn()
            finally:
                ex = ctypes.c_void_p()
                status = _lib.load().mbem_graph_end(stream(), ctypes.byref(ex))
            if status != _lib.MBEM_OK:
                raise _lib.MbemError(f"HIP graph capture of {key} failed ({status})")
            self._graphs[key] = ex
        call("mbem_graph_launch", self._graphs[key], stream())

    def release(self):
        for ex in self._graphs.values():
            _lib.load().mbem_graph_destroy(ex)
        self._graphs.clear()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # -- solve -------------------------------------------------------------------
    def solve(self, B: torch.Tensor, rtol: float = 1e-10, maxcycles: int = 20) -> BatchedReport:
        if B.shape != (self.n, self.K):
            raise ValueError(f"right-hand side must be {(self.n, self.K)}")
        if self.graph and torch.cuda.current_stream() == torch.cuda.default_stream():
            raise _lib.MbemError("graph capture needs a non-default stream (use torch.cuda.stream)")
        t0 = time.perf_counter()
        st = torch.cuda.current_stream()
        self.M._table(self.M.cfg)           # device work table, made before any capture
        self.op._prepare()
        self.B.copy_(B)
        bnorm = torch.linalg.vector_norm(self.B, dim=0).cpu().numpy()
        self.X.zero_()
        call("mbem_bcopy", ptr(self.B), ptr(self.R), self.n * self.K, stream())
        self._start()           # eager: the first start is not part of a cycle graph
        st.synchronize()
        tol = rtol * bnorm
        iters = np.full(self.K, -1, dtype=np.int64)
        steps = 0
        cycles = 0
        beta = self.beta_host.numpy().copy()
        done = beta <= tol
        iters[done] = 0
        nseg = (self.m + self.segment - 1) // self.segment
        while not done.all() and cycles < maxcycles:
            cycles += 1
            for q in range(nseg):
                self._run(("seg", q), lambda q=q: self._segment(q))
                st.synchronize()
                steps += min(self.segment, self.m - q * self.segment)
                est = self.res_host.numpy()
                newly = (iters < 0) & (est <= tol)
                iters[newly] = steps
                if np.all((est <= tol) | done):
                    break
            # the end graph solves over the steps done this cycle (read on the device)
            self._run(("end",), self._end)
            st.synchronize()
            beta = self.beta_host.numpy().copy()
            done = beta <= tol
        relres = np.where(bnorm > 0, beta / np.where(bnorm > 0, bnorm, 1.0), 0.0)
        iters[(iters < 0) & done] = steps
        return BatchedReport(converged=done, iterations=iters, cycles=cycles, relres=relres,
                             seconds=time.perf_counter() - t0, graph=self.graph, steps=steps)
