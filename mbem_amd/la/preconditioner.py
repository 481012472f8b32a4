"""Block Gauss-Seidel right preconditioner over super-blocks, on the device
(mbem's src/mbem/la/preconditioner.py:77-128, :314-543).

Grouping: ``"patch"`` -- one super-block per slot, an interface's (u, t)
pair together; ``"region"`` -- one per region (group_slots, :77-112).  One
forward Gauss-Seidel sweep in block order: block k solves its diagonal block
against ``r_k - A[k, <k] z_<k]`` (:521-543).  Diagonal solves form the
reference's ladder: a block of at most ``max_dense`` DOFs is solved whole
("dense_lu" rung), a larger one by cluster block-Jacobi -- the median-
bisection cluster tree over the block's collocation points cut into chunks of
at most ``bj_chunk`` DOFs (at least 32 units), each chunk solved exactly
(:421-470).  The chunk matrices are gathered from the assembled operator
itself (what ``_SBEvaluator.eval_block`` re-evaluates, :230-292; the gate the
reference keeps says the two agree entry for entry).

MI355X form: each chunk is factored by LAPACK getrf (rocSOLVER) and kept as
its explicit inverse (getri), so one application of the whole ladder is a
batched GEMV over the inverses (HBM-bound, one launch per super-block) plus
one GEMV for the lower couplings, instead of two dependent triangular solves
per chunk.  Same preconditioner in exact arithmetic; in floating point the
inverse differs from the LU solve by cond(chunk) * eps, which a flexible
outer iteration absorbs (the outer residual is the true one).
"""
from __future__ import annotations

import ctypes
import time
import weakref

import numpy as np
import torch

from .. import _lib, defaults
from .._lib import call, ptr, stream
from ..model.core import BCType
from .cluster import build_cluster_tree, leaves

F64 = torch.float64
I64 = torch.int64
# largest FP32 chunk the batched Gauss-Jordan inverts (four row slots of its
# 512-thread workgroup); larger chunks go to rocSOLVER
BJ_GJ_MAX_DOF = 2048


def group_slots(model, layout, grouping: str) -> list:
    """Super-blocks as lists of slots in sweep order (preconditioner.py:77-112)."""
    if grouping == "patch":
        groups, seen = [], set()
        for slot in layout.slots:
            if id(slot.patch) in seen:
                continue
            if slot.patch.bc is BCType.INTERFACE:
                seen.add(id(slot.patch))
                groups.append([layout.slot(slot.patch, "u"), layout.slot(slot.patch, "t")])
            else:
                groups.append([slot])
        return groups
    if grouping == "region":
        owner = {}
        for k, region in enumerate(model.regions):
            for patch in region.patches:
                owner[id(patch)] = k
        groups = []
        for k in range(len(model.regions)):
            slots = [s for s in layout.slots if owner[id(s.patch)] == k]
            if slots:
                groups.append(slots)
        return groups
    raise ValueError(f"unknown preconditioner grouping {grouping!r} (expected 'patch' or "
                     f"'region')")


def dense_rung_max_dof(max_dense: int = defaults.MAX_DENSE_PRECOND_DOF,
                       mem_fraction: float = defaults.PRECOND_DENSE_RAM_FRACTION) -> int:
    """Largest block the whole-block rung takes (preconditioner.py:115-128),
    with the device's memory in the role of the host RAM."""
    try:
        mem = torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory
    except Exception:
        return max_dense
    return min(max_dense, int(np.sqrt(mem_fraction * mem / 8.0)))


class _Units:
    """Unit layout of one super-block (preconditioner.py:139-228): one unit
    per (slot group, element), d = 3 x slots per group; groups of unequal
    slot counts fall back to one group per slot."""

    def __init__(self, slots, points_of):
        by_patch, order = {}, []
        for s in slots:
            if id(s.patch) not in by_patch:
                by_patch[id(s.patch)] = []
                order.append(id(s.patch))
            by_patch[id(s.patch)].append(s)
        groups = [by_patch[k] for k in order]
        if len({len(g) for g in groups}) > 1:
            groups = [[s] for s in slots]
        self.groups = groups
        self.d = 3 * len(groups[0])
        pts, gid, eid = [], [], []
        for gi, g in enumerate(groups):
            p = points_of(g[0].patch)
            pts.append(p)
            gid.append(np.full(p.shape[0], gi))
            eid.append(np.arange(p.shape[0]))
        self.centroids = np.ascontiguousarray(np.concatenate(pts))
        self.group_id = np.concatenate(gid)
        self.elem_id = np.concatenate(eid)
        self.n_units = self.centroids.shape[0]

    def dofs(self, units) -> np.ndarray:
        """Global DOF indices of ``units``, unit-major: d per unit."""
        out = np.empty((len(units), self.d), dtype=np.int64)
        for gi, g in enumerate(self.groups):
            sel = np.nonzero(self.group_id[units] == gi)[0]
            e = self.elem_id[units][sel]
            for k, s in enumerate(g):
                for a in range(3):
                    out[sel, 3 * k + a] = s.offset + 3 * e + a
        return out.ravel()


class _SuperBlock:
    def __init__(self, slots):
        self.slots = list(slots)
        self.start = min(s.offset for s in self.slots)
        self.stop = max(s.stop for s in self.slots)
        self.size = sum(s.size for s in self.slots)
        if self.stop - self.start != self.size:
            raise NotImplementedError("a super-block whose slots are not one contiguous range "
                                      "of unknowns is not supported")
        self.rung = None
        self.build_s = 0.0
        self.detail: dict = {}
        self.wave_table = None
        self.n_waves = 0


class BlockGaussSeidel:
    """z = M(r) for an assembled dense operator (preconditioner.py:314-543).

    ``max_dense``: None = the device's own cap; 0 forces block-Jacobi on every
    super-block (what the reference's gates do to force a rung).
    ``bj_chunk``: chunk DOF cap (default PRECOND_BJ_CHUNK_DOF)."""

    def __init__(self, assembled, max_dense=None, bj_chunk=None,
                 grouping: str = defaults.PRECOND_GROUPING,
                 above_dense: str = defaults.PRECOND_RUNG_ABOVE_DENSE, verbose: bool = False,
                 hodlr_tol=None, hodlr_max=None, inverse_dtype: str = "f64", coupling=None,
                 defer_check: bool = False, after_gather=None):
        _lib.require_gpu()
        if max_dense is None:
            max_dense = dense_rung_max_dof()
        if bj_chunk is None:
            bj_chunk = defaults.PRECOND_BJ_CHUNK_DOF
        if above_dense != "block_jacobi":
            raise NotImplementedError(f"rung {above_dense!r} above the dense cap: only "
                                      f"'block_jacobi' is in mbem_amd")
        if grouping == "region" and getattr(assembled, "deflate", False):
            grouping = "patch"
        if inverse_dtype not in ("f64", "f32"):
            raise ValueError(f"inverse_dtype {inverse_dtype!r} (expected 'f64' or 'f32')")
        self.inverse_dtype = inverse_dtype
        self.coupling = coupling
        self.max_dense = int(max_dense)
        self.bj_chunk = int(bj_chunk)
        self.above_dense = above_dense
        self.grouping = grouping
        # the assembly is held weakly: it owns this preconditioner
        # (``_precond``), so a strong reference would make a cycle that keeps
        # its n x n matrix alive after the caller drops it
        self._asm = weakref.ref(assembled)
        self.n_unknowns = assembled.layout.n_unknowns
        A = assembled.A_dev
        self.A = A
        n = A.shape[0]
        dev = A.device
        layout = assembled.layout
        self.sbs = [_SuperBlock(g) for g in group_slots(assembled.system.model, layout, grouping)]
        self.sbs.sort(key=lambda sb: sb.start)

        t0 = time.perf_counter()
        chunks = []                    # (sb index, dof array)
        for k, sb in enumerate(self.sbs):
            t_sb = time.perf_counter()
            units = _Units(sb.slots, assembled.collocation_points_of)
            if sb.size <= self.max_dense:
                parts = [np.arange(units.n_units)]
                sb.rung = "dense_lu"
            else:
                chunk_units = max(32, self.bj_chunk // units.d)
                parts = leaves(build_cluster_tree(units.centroids, chunk_units))
                sb.rung = "block_jacobi"
                sb.detail = {"chunks": len(parts)}
            for p in parts:
                chunks.append((k, units.dofs(np.asarray(p))))
            sb.build_s = time.perf_counter() - t_sb
            if verbose:
                print(f"  SB {[s.name for s in sb.slots]}: {sb.rung} ({sb.size} DOFs, "
                      f"{len(parts)} chunk(s))")
        # device tables: chunk DOF lists, inverse offsets, per-block wave tables
        sizes = np.array([c[1].size for c in chunks], dtype=np.int64)
        cptr = np.concatenate([[0], np.cumsum(sizes)])
        f32 = inverse_dtype == "f32"
        # FP32 inverses keep their rows padded to a multiple of 4 floats
        # (16-byte loads; pads zero) and gather r once per apply into the
        # same padded chunk order (gidx: the DOF of each padded slot, -1 for
        # a pad); FP64 inverses are packed m x m
        ld = ((sizes + 3) // 4) * 4 if f32 else sizes
        iptr = np.concatenate([[0], np.cumsum(sizes * ld)])
        # pinned, non-blocking uploads: the host goes on to queue the
        # inversions behind whatever the stream still runs (the assembly)
        self.chunk_ptr = _lib.upload(cptr, I64)
        self.inv_ptr = _lib.upload(iptr, I64)
        self.idx = _lib.upload(np.concatenate([c[1] for c in chunks]), I64)
        # FP32 chunks up to BJ_GJ_MAX_DOF are inverted by the project's batched
        # Gauss-Jordan (every element of the packed inverse written), larger
        # ones by rocSOLVER per size class
        use_gj = f32 and sizes.size > 0 and int(sizes.max()) <= BJ_GJ_MAX_DOF
        if f32:
            pptr = np.concatenate([[0], np.cumsum(ld)])
            gidx = np.full(int(pptr[-1]), -1, dtype=np.int64)
            for ci, (_, dofs) in enumerate(chunks):
                gidx[pptr[ci]:pptr[ci] + dofs.size] = dofs
            self.pad_ptr = _lib.upload(pptr, I64)
            self.gidx = _lib.upload(gidx, I64)
            self.rr32 = torch.zeros(int(pptr[-1]), dtype=torch.float32, device=dev)
            self.inv = (torch.empty if use_gj else torch.zeros)(int(iptr[-1]),
                                                                dtype=torch.float32, device=dev)
        else:
            self.inv = torch.empty(int(iptr[-1]), dtype=F64, device=dev)
        self._ld = ld
        R = _lib.BJ32_ROWS_PER_WAVE if f32 else _lib.BJ_ROWS_PER_WAVE
        tables = {k: [] for k in range(len(self.sbs))}
        first = {}
        for ci, (k, _) in enumerate(chunks):
            first.setdefault(k, ci)
        for k, sb in enumerate(self.sbs):
            c0 = first.get(k, len(chunks))
            c1 = c0 + sum(1 for c in chunks if c[0] == k)
            sb.chunks = (c0, c1)
            if f32:
                sb.pad = (int(pptr[c0]), int(pptr[c1]))
        for ci, (k, dofs) in enumerate(chunks):
            tables[k] += [(ci, r0) for r0 in range(0, int(dofs.size), R)]
        if use_gj:
            self._pending = self._invert_gj(A, chunks, sizes, cptr, iptr, ld, after_gather)
        else:
            # invert the chunks in size classes (largest / smallest <= 1.1): each
            # class is padded to its largest size with an identity block and
            # inverted in one strided-batched getrf + getri, then the leading
            # blocks are packed into ``inv``.  Every class is gathered first, and
            # ``gathered`` (an event on the stream) marks the last read of A, so a
            # caller may change A on another stream after it while the inversions
            # run on the padded copies.
            order = np.argsort(sizes, kind="stable")
            esz = 4 if f32 else 8
            classes = []
            c0 = 0
            while c0 < len(order):
                c1 = c0
                while c1 + 1 < len(order) and sizes[order[c1 + 1]] <= 1.1 * sizes[order[c0]]:
                    c1 += 1
                cls = order[c0:c1 + 1]
                mp = int(sizes[cls].max())
                tmp = torch.empty((len(cls), mp, mp), dtype=torch.float32 if f32 else F64,
                                  device=dev)
                for q, ci in enumerate(cls):
                    call("mbem_gather_block_pad", ptr(A), A.stride(0),
                         ptr(self.idx[int(cptr[ci]):int(cptr[ci + 1])]), int(sizes[ci]), mp,
                         ptr(tmp[q]), int(f32), stream())
                classes.append((cls, mp, tmp))
                c0 = c1 + 1
            # host copies of the chunk DOF lists (a coarse space is built on them;
            # ``after_gather`` receives them with the event)
            self.chunk_dofs = [c[1] for c in chunks]
            self.gathered = torch.cuda.Event()
            self.gathered.record()
            if after_gather is not None:
                # the caller's own launches (e.g. on another stream, after
                # ``gathered``) go ahead of the inversions' long launch sequence
                after_gather(self.gathered, self.chunk_dofs)
            # getrf's info flags, read back once after every class is queued
            # (``finish``), so no class waits on the one before.  The inverse of
            # each class goes to a second buffer (getrf, then A C = I solved with
            # the factors), then its leading blocks into ``inv``.
            self._pending = []
            for cls, mp, tmp in classes:
                nb = len(cls)
                ipiv = torch.empty(nb * mp, dtype=torch.int32, device=dev)
                info = torch.zeros(nb, dtype=torch.int32, device=dev)
                out = torch.empty_like(tmp)
                call("mbem_invert_batched_oop", ptr(tmp), mp, nb, ptr(ipiv), ptr(info), ptr(out),
                     int(f32), stream())
                self._pending.append((info, cls, sizes[cls]))
                for q, ci in enumerate(cls):
                    m = int(sizes[ci])
                    call("mbem_copy2d", ptr(out[q]), mp, ptr(self.inv[int(iptr[ci]):]),
                         int(ld[ci]), m, m, esz, stream())
            del classes
        for k, sb in enumerate(self.sbs):
            wt = np.asarray(tables[k], dtype=np.int64).reshape(-1, 2)
            sb.wave_table = _lib.upload(wt, I64)
            sb.n_waves = wt.shape[0]
        self._t0 = t0
        self.build_s = None
        if not defer_check:
            self.finish()
        self.n_chunks = len(chunks)
        self.inverse_bytes = int(iptr[-1]) * (4 if f32 else 8)     # padding included
        self._rr = torch.empty(n, dtype=F64, device=dev)

    def _invert_gj(self, A, chunks, sizes, cptr, iptr, ld, after_gather) -> list:
        """Every chunk's FP32 inverse into ``inv`` in one launch sequence: the
        chunks gathered from A (then ``gathered``, as for the size classes),
        the batched blocked Gauss-Jordan of mbem_bj_invert_f32 queued
        (partial pivoting by a row map, 32-column panels, two launches per
        panel for all chunks), then the ``after_gather`` hook.  Returns the
        pending status flags for ``finish``."""
        lib = _lib.load()
        nf, B = int(lib.mbem_bj_desc_fields()), int(lib.mbem_bj_panel_width())
        nc = int(sizes.size)
        mmax = int(sizes.max())
        ldw = ((sizes + B - 1) // B) * B
        woff = np.concatenate([[0], np.cumsum(sizes * ldw)])
        soff = np.concatenate([[0], np.cumsum(B * ldw)])
        desc = np.zeros((nc, nf), dtype=np.int64)
        desc[:, 0] = cptr[:-1]          # the chunk's DOFs in idx
        desc[:, 1] = sizes
        desc[:, 2] = woff[:-1]          # working copy, pitch a multiple of the panel width
        desc[:, 3] = ldw
        desc[:, 4] = iptr[:-1]          # the packed inverse (pitch ld)
        desc[:, 5] = ld
        desc[:, 6] = cptr[:-1]          # row map and pivots
        desc[:, 7] = soff[:-1]          # saved pivot rows
        dev = A.device
        d_dev = _lib.upload(desc, I64)
        W = torch.empty(int(woff[-1]), dtype=torch.float32, device=dev)
        rowmap = torch.empty(int(cptr[-1]), dtype=torch.int32, device=dev)
        piv = torch.empty_like(rowmap)
        saved = torch.empty(int(soff[-1]), dtype=torch.float32, device=dev)
        info = torch.empty(nc, dtype=torch.int32, device=dev)
        dh = ctypes.c_void_p(desc.ctypes.data)
        call("mbem_bj_gather_f32", ptr(A), A.stride(0), ptr(self.idx), int(cptr[-1]), dh,
             ptr(d_dev), nc, mmax, ptr(W), W.numel(), ptr(rowmap), rowmap.numel(), ptr(info),
             stream())
        self.chunk_dofs = [c[1] for c in chunks]
        self.gathered = torch.cuda.Event()
        self.gathered.record()
        # the inversion's few launches go ahead of the hook's host work (which
        # waits for the near pattern and queues the caller's side-stream build)
        call("mbem_bj_invert_f32", dh, ptr(d_dev), nc, mmax, ptr(W), W.numel(), ptr(rowmap),
             ptr(piv), rowmap.numel(), ptr(saved), saved.numel(), ptr(info), ptr(self.inv),
             self.inv.numel(), stream())
        if after_gather is not None:
            after_gather(self.gathered, self.chunk_dofs)
        # the working memory stays referenced until ``finish`` has waited
        self._gj_work = (W, rowmap, piv, saved, d_dev)
        return [(info, np.arange(nc), sizes)]

    def finish(self) -> "BlockGaussSeidel":
        """Wait for the chunk inversions and raise ``LinAlgError`` for the first
        singular chunk (getrf's info), as the build would have at once.  Called
        by the constructor unless ``defer_check``; a caller that defers it
        queues more device work first and calls this before the first apply.
        ``build_s`` is then the time from the build's start to here."""
        if self._pending:
            flags = torch.cat([p[0] for p in self._pending]).cpu().numpy()
            o = 0
            for info, cls, sz in self._pending:
                bad = flags[o:o + cls.size]
                o += cls.size
                if np.any(bad != 0):
                    q = int(np.flatnonzero(bad)[0])
                    raise np.linalg.LinAlgError(f"block-Jacobi chunk {int(cls[q])} "
                                                f"({int(sz[q])} DOFs) is singular "
                                                f"(zero pivot, info {int(bad[q])})")
            self._pending = []
        torch.cuda.current_stream().synchronize()
        self._gj_work = None
        if self.build_s is None:
            self.build_s = time.perf_counter() - self._t0
        return self

    def apply_into(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        """z = M(r): one forward sweep (preconditioner.py:521-543)."""
        A, rr = self.A, self._rr
        for sb in self.sbs:
            seg = slice(sb.start, sb.stop)
            if sb.start == 0:
                # the first super-block has no lower coupling: read r itself
                self._apply_block(sb, r, z)
                continue
            rr[seg].copy_(r[seg])
            if sb.start > 0:
                if self.coupling is not None:
                    self.coupling.lower_into(sb.start, sb.stop, z, rr[seg])
                else:
                    call("mbem_gemv", ptr(A[sb.start:sb.stop]), A.stride(0), sb.size, sb.start,
                         ptr(z), ptr(rr[seg]), -1.0, 1.0, stream())
            self._apply_block(sb, rr, z)
        return z

    def _apply_block(self, sb, r, z):
        """z[sb's DOFs] = (its chunk inverses) r[sb's DOFs]."""
        if self.inverse_dtype == "f32":
            p0, p1 = sb.pad
            call("mbem_gather_f32", ptr(self.gidx[p0:p1]), p1 - p0, ptr(r), ptr(self.rr32[p0:p1]),
                 stream())
            call("mbem_bj_apply_f32v", ptr(sb.wave_table), sb.n_waves, ptr(self.chunk_ptr),
                 ptr(self.pad_ptr), ptr(self.idx), ptr(self.inv_ptr), ptr(self.inv),
                 ptr(self.rr32), ptr(z), stream())
        else:
            call("mbem_bj_apply", ptr(sb.wave_table), sb.n_waves, ptr(self.chunk_ptr),
                 ptr(self.idx), ptr(self.inv_ptr), ptr(self.inv), ptr(r), ptr(z), stream())

    def __call__(self, r: torch.Tensor) -> torch.Tensor:
        z = torch.empty_like(r)
        return self.apply_into(r, z)

    def inverse_diagonal(self) -> torch.Tensor:
        """diag(M) as an FP64 device vector over the unknowns: the diagonal of
        each chunk inverse scattered to its DOFs."""
        n = self.n_unknowns
        d = torch.zeros(n, dtype=F64, device=self.inv.device)
        cptr = self.chunk_ptr.cpu().numpy()
        iptr = self.inv_ptr.cpu().numpy()
        for c in range(len(cptr) - 1):
            m = int(cptr[c + 1] - cptr[c])
            ld = int(self._ld[c])
            M = self.inv[int(iptr[c]):int(iptr[c]) + m * ld].view(m, ld)[:, :m]
            d[self.idx[int(cptr[c]):int(cptr[c + 1])]] = M.diagonal().to(F64)
        return d

    @property
    def asm(self):
        """The assembly this preconditioner was built from (None once it is
        gone)."""
        return self._asm()

    def summary(self) -> dict:
        """preconditioner.py:502-519."""
        return {
            "grouping": self.grouping, "above_dense": self.above_dense,
            "max_dense": self.max_dense, "bj_chunk": self.bj_chunk, "build_s": self.build_s,
            "chunks": self.n_chunks, "inverse_bytes": self.inverse_bytes,
            "inverse_dtype": self.inverse_dtype,
            "super_blocks": [{"slots": [s.name for s in sb.slots], "size": sb.size,
                              "rung": sb.rung, "build_s": sb.build_s, **sb.detail}
                             for sb in self.sbs],
        }
