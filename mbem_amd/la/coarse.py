"""Two-level block-Jacobi: the chunk inverses plus a coarse correction.

Block-Jacobi (la/preconditioner.py) leaves the error that varies slowly
from chunk to chunk; a coarse space W spanned by a few shape functions on
every chunk takes it out.  Each group -- one block-Jacobi chunk's DOFs of
one slot kind (displacement or traction unknowns) -- carries q functions of
the group's own geometry (positions p of its collocation points, centred on
the group and scaled to unit RMS radius):

* ``"const"``  (q = 3): the three constant vector fields e_b;
* ``"rigid"``  (q = 6): those and the three rotations p x e_b;
* ``"linear"`` (q = 12): the constants and the nine fields p_e e_b.

The correction is the deflated ("coarse first") form.  With C = A W and Q
an orthonormal basis of range(C) (C T = Q):

    s = Q^T r,   z = M (r - Q s) + W T s,

so A z = Q s + A M (r - Q s): the part of r in range(A W) is solved
exactly and block-Jacobi sees only the rest.  For r = A W y, z = W y
(to the orthogonality of Q).  The build is one pass over A
(``mbem_coarse_image``: Ct = C^T), the Gram matrix C^T C on matrix cores,
its regularised Cholesky factor and inverse (rocSOLVER, K x K) and
Q^T = T^T Ct (one GEMM), all queued without a host wait; each apply is a
multi-dot and a multi-axpy over Q (8 n K bytes each) plus the expansion of
W T s.

This is an addition to the reference's preconditioner ladder
(la/preconditioner.py:421-470 has block-Jacobi without a coarse level).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from .._lib import call, ptr, stream

F64 = torch.float64
I64 = torch.int64
BASES = {"const": 3, "rigid": 6, "linear": 12}


def dof_geometry(assembled):
    """(point (n, 3), kind (n,) 0 = u / 1 = t, component (n,)) of every
    unknown: the collocation point of its element."""
    layout = assembled.layout
    n = layout.n_unknowns
    pts = np.empty((n, 3))
    kind = np.empty(n, np.int64)
    comp = np.tile(np.arange(3), n // 3)
    for s in layout.slots:
        x = assembled.collocation_points_of(s.patch)
        pts[s.offset:s.stop] = np.repeat(x, 3, axis=0)
        kind[s.offset:s.stop] = 0 if s.kind == "u" else 1
    return pts, kind, comp


# (p x e_b)_a = sum_j eps_{a j b} p_j: for a != b the one nonzero term has
# j = 3 - a - b and the sign of the permutation (a, j, b)
_ROT_J = np.array([[(3 - a - b) % 3 if a != b else 0 for b in range(3)] for a in range(3)])
_ROT_S = np.array([[0.0 if a == b else float(np.linalg.det(np.eye(3)[[a, 3 - a - b, b]]))
                    for b in range(3)] for a in range(3)])


def shape_values(p, comp, basis: str) -> np.ndarray:
    """(len(p), q) values of the group's q shape functions at DOFs with
    positions p (already centred and scaled) and components comp."""
    q = BASES[basis]
    m = p.shape[0]
    out = np.zeros((m, q))
    rows = np.arange(m)
    out[rows, comp] = 1.0                                   # constants e_b
    if basis == "rigid":
        for b in range(3):                                  # rotations p x e_b
            out[:, 3 + b] = _ROT_S[comp, b] * p[rows, _ROT_J[comp, b]]
    elif basis == "linear":
        for e in range(3):                                  # p_e e_b
            out[rows, 3 + 3 * e + comp] = p[:, e]
    return out


def coarse_groups(chunk_dofs, pts, kind, comp, basis: str):
    """(perm, gptr, Wp): the DOFs of every (chunk, kind) group listed group
    by group (chunks in order, displacement before traction DOFs, ascending
    DOFs within a group: the image kernel's gathered row reads then walk
    each row forward in runs), and the shape values there."""
    parts, gptr = [], [0]
    for dofs in chunk_dofs:
        d = np.sort(np.asarray(dofs, np.int64))
        k = kind[d]
        for kk in (0, 1):
            g = d[k == kk] if k.min() != k.max() else (d if k[0] == kk else d[:0])
            if g.size:
                parts.append(g)
                gptr.append(gptr[-1] + g.size)
    perm = np.concatenate(parts)
    gptr = np.asarray(gptr, np.int64)
    cnt = np.diff(gptr)
    gid = np.repeat(np.arange(cnt.size), cnt)
    p = pts[perm]
    mean = np.add.reduceat(p, gptr[:-1], axis=0) / cnt[:, None]
    p = p - mean[gid]
    rms = np.sqrt(np.add.reduceat(np.sum(p * p, axis=1), gptr[:-1]) / cnt)
    p = p / np.where(rms > 0.0, rms, 1.0)[gid][:, None]
    return perm, gptr, np.ascontiguousarray(shape_values(p, comp[perm], basis))


class CoarseSpace:
    """W, Q = orth(A W) and T (A W T = Q) on the device, for the operator
    ``A`` (an n x n row-major FP64 device matrix, ``assembled.A_dev`` by
    default) and the DOF chunks ``chunk_dofs`` of a block-Jacobi built on the
    same unknowns.

    In two phases, neither waiting for the device: the constructor uploads W
    (pinned, non-blocking) and queues the coarse image Ct = (A W)^T and its
    Gram matrix Ct Ct^T on the current stream; ``finish()`` (called by the
    first apply if not before) queues the factorisation and Q^T.  A caller
    can queue the first phase on a side stream beside other work.
    ``op``, a MixedOperator with int8 storage, takes the image of that
    operator (its int8 far field and FP64 near blocks) instead of A's."""

    def __init__(self, assembled, chunk_dofs, basis: str = "rigid", A: torch.Tensor = None,
                 drop: float = 1e-13, op=None, extra: torch.Tensor = None):
        if basis not in BASES and basis != "none":
            raise ValueError(f"coarse basis {basis!r} (expected one of {sorted(BASES)} or 'none')")
        _lib.require_gpu()
        A = assembled.A_dev if A is None else A
        n = A.shape[0]
        self.n, self.basis, self.drop = n, basis, drop
        self.q = BASES.get(basis, 1)
        pts, kind, comp = dof_geometry(assembled)
        perm, gptr, Wp = coarse_groups(chunk_dofs, pts, kind, comp,
                                       basis if basis != "none" else "const")
        if perm.size != n or np.any(np.bincount(perm, minlength=n) != 1):
            raise ValueError("the block-Jacobi chunks do not cover every unknown exactly once")
        ngroups = gptr.size - 1
        self.G = ngroups if basis != "none" else 0
        # extra: m global columns (n x m, row-major FP64, the unknowns' own
        # order) beside the groups' shape functions
        self.extra = None if extra is None else extra.to(F64).contiguous()
        if self.extra is not None and (self.extra.dim() != 2 or self.extra.shape[0] != n):
            raise ValueError(f"extra coarse columns must be (n, m) = ({n}, m)")
        self.m = 0 if self.extra is None else self.extra.shape[1]
        self.K_local = self.G * self.q
        K = self.K_local + self.m
        # the image and projection of an int8 operator run in FP32 (below)
        self.q32 = op is not None and getattr(op, "i8", False) and self.K_local > 0
        self.K = K
        self.perm = _lib.upload(perm, I64)
        self.Wp = _lib.upload(Wp, F64)
        grp = np.repeat(np.arange(ngroups, dtype=np.int64), np.diff(gptr))
        self.grp = _lib.upload(grp, I64)
        self._gptr = _lib.upload(gptr, I64)
        dev = A.device
        self._Ct = torch.empty((K, n), dtype=F64, device=dev)
        if self.m:
            # the image of the global columns: one library GEMM A E, then its
            # transpose into the last m rows of Ct
            AE = torch.empty((n, self.m), dtype=F64, device=dev)
            call("mbem_dgemm_rowmajor", n, self.m, n, 1.0, ptr(A), A.stride(0), ptr(self.extra),
                 self.m, 0.0, ptr(AE), self.m, stream())
            self._Ct[self.K_local:].copy_(AE.t())
            del AE
        if self.K_local == 0:
            self.source = "FP64 matrix (global columns only)"
        elif self.q32:
            # the image of the mixed operator F + N itself: the int8 far
            # field (n^2 bytes) and the FP64 near blocks
            inv = np.empty(n, np.int64)
            inv[perm] = np.arange(n, dtype=np.int64)
            self._inv = _lib.upload(inv, I64)
            # row chunk by row chunk behind the operator's cast events (when
            # this runs on another stream than the cast, the image of the
            # first chunks overlaps the cast of the later ones)
            chunks = getattr(op, "cast_events", None) or [(0, n, None)]
            cur = torch.cuda.current_stream()
            for r0, r1, ev in chunks:
                if ev is not None:
                    cur.wait_event(ev)
                call("mbem_coarse_image_i8", ptr(op.F[r0:r1]), op.ldf, ptr(op.S[r0:r1]), op.nseg,
                     r1 - r0, ptr(self.perm), ptr(self.Wp), ptr(self._gptr), self.G, self.q,
                     ptr(self._Ct[:, r0:]), n, stream())
            call("mbem_coarse_image_bsr", op.nbr, ptr(op.rowptr), ptr(op.col), ptr(op.vals),
                 ptr(self._inv), ptr(self.grp), ptr(self.Wp), self.q, ptr(self._Ct), n, stream())
            self.source = "int8 far field + FP64 near field"
        else:
            call("mbem_coarse_image", ptr(A), A.stride(0), n, ptr(self.perm), ptr(self.Wp),
                 ptr(self._gptr), self.G, self.q, ptr(self._Ct), n, stream())
            self.source = "FP64 matrix"
        self._gram = torch.empty((K, K), dtype=F64, device=dev)
        gwork = torch.empty(int(_lib.load().mbem_gram_work(K, n)), dtype=F64, device=dev)
        call("mbem_gram", ptr(self._Ct), n, K, n, ptr(self._gram), ptr(gwork), stream())
        gwork.record_stream(torch.cuda.current_stream())
        self.Qt = None
        self.bytes = 0

    def finish(self) -> "CoarseSpace":
        """Queue the factorisation behind the image (no host wait): the
        Gram matrix scaled to unit diagonal (displacement and traction
        columns differ by ~1e12) plus ``drop`` I, its Cholesky factor and
        inverse, T, and Q^T = T^T Ct with one GEMM.  Built from an int8
        operator (``op``), Q^T is kept in FP32 for the projections (FP64
        vectors and sums): that operator is A to 1.8e-4, FP32 changes
        nothing it does, and it halves the 2 K n bytes a projection reads;
        otherwise Q^T stays FP64.  The regularisation keeps every direction
        of the scaled Gram above ``drop`` and fades out the
        (near-)dependent ones, so no pivoting or rank decision needs the
        host."""
        if self.Qt is not None:
            return self
        K, n, dev = self.K, self.n, self._Ct.device
        d = torch.empty(max(K, 1), dtype=F64, device=dev)
        info = torch.zeros(2, dtype=torch.int32, device=dev)
        Tt = torch.empty((K, K), dtype=F64, device=dev)
        self.T = torch.empty((K, K), dtype=F64, device=dev)
        call("mbem_coarse_factor", ptr(self._gram), K, float(self.drop), ptr(d), ptr(info),
             ptr(Tt), ptr(self.T), stream())
        self.info = info
        # Q^T in FP64 (row pitch ldq, 16-byte rows); for the FP32 projection
        # its FP32 copy
        self.ldq = ((n + 3) // 4) * 4 if self.q32 else n
        q64 = torch.empty((K, self.ldq), dtype=F64, device=dev)
        call("mbem_dgemm_rowmajor", K, n, K, 1.0, ptr(Tt), K, ptr(self._Ct), n, 0.0,
             ptr(q64), self.ldq, stream())
        if self.q32:
            self.Qt = torch.empty((K, self.ldq), dtype=torch.float32, device=dev)
            call("mbem_vec_to_f32", ptr(q64), K * self.ldq, ptr(self.Qt), stream())
            q64.record_stream(torch.cuda.current_stream())
        else:
            self.Qt = q64
        # the image may have been built on another stream: its blocks go back
        # to the allocator only after this stream's GEMM has read them
        for t in (self._Ct, self._gram):
            t.record_stream(torch.cuda.current_stream())
        self._Ct = self._gram = None
        del q64
        self.Kp = K
        self.s = torch.empty(max(K, 1), dtype=F64, device=dev)
        self.u = torch.empty(max(K, 1), dtype=F64, device=dev)
        lib = _lib.load()
        work = (lib.mbem_coarse_project_work(max(K, 1), n) if self.q32 else
                lib.mbem_multidot_work(max(K, 1), n))
        self.work = torch.empty(int(work), dtype=F64, device=dev)
        self.bytes = (K * self.ldq * (4 if self.q32 else 8) + K * K * 8 + self.Wp.numel() * 8
                      + 2 * n * 8 + 8 * self.m * n)
        return self

    def check(self):
        """Raise if the factorisation failed (reads its two info flags)."""
        flags = self.finish().info.cpu().numpy()
        if np.any(flags != 0):
            raise np.linalg.LinAlgError(f"coarse Gram factorisation: potrf / trtri info {flags}")

    def project(self, r: torch.Tensor, rest: torch.Tensor) -> torch.Tensor:
        """s = Q^T r (kept in ``self.s``) and rest = r - Q s."""
        self.finish()
        Kp, n = self.Kp, self.n
        if Kp == 0:
            rest.copy_(r)
            return rest
        if self.q32:
            call("mbem_coarse_project_f32", ptr(self.Qt), self.ldq, Kp, n, ptr(r), ptr(self.s),
                 ptr(rest), ptr(self.work), stream())
            return rest
        rest.copy_(r)
        call("mbem_multidot", ptr(self.Qt), n, Kp, n, ptr(r), ptr(self.s), ptr(self.work), stream())
        call("mbem_multiaxpy", ptr(self.Qt), n, Kp, n, ptr(self.s), ptr(rest), -1.0, stream())
        return rest

    def add_coarse(self, z: torch.Tensor) -> torch.Tensor:
        """z += W T s for the ``s`` of the last ``project``."""
        if self.Kp == 0:
            return z
        call("mbem_gemv", ptr(self.T), self.Kp, self.K, self.Kp, ptr(self.s), ptr(self.u), 1.0, 0.0,
             stream())
        if self.K_local:
            call("mbem_coarse_expand", ptr(self.u), ptr(self.perm), ptr(self.Wp), ptr(self.grp),
                 self.n, self.q, 1.0, ptr(z), stream())
        if self.m:
            call("mbem_gemv", ptr(self.extra), self.m, self.n, self.m, ptr(self.u[self.K_local:]),
                 ptr(z), 1.0, 1.0, stream())
        return z


class TwoLevel:
    """z = P(r) = M (r - Q Q^T r) + W T Q^T r: block-Jacobi (or the block
    Gauss-Seidel sweep) ``M`` with the coarse correction of ``coarse``."""

    def __init__(self, M, coarse: CoarseSpace):
        self.M, self.coarse = M, coarse
        self._rest = torch.empty(coarse.n, dtype=F64, device=coarse.Wp.device)

    def apply_into(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        rest = self.coarse.project(r, self._rest)
        self.M.apply_into(rest, z)
        return self.coarse.add_coarse(z)

    def __call__(self, r):
        z = torch.empty_like(r)
        return self.apply_into(r, z)

    def summary(self) -> dict:
        s = self.M.summary()
        c = self.coarse.finish()
        s["coarse"] = {"basis": c.basis, "groups": c.G, "columns": c.K, "global_columns": c.m,
                       "rank": c.Kp, "bytes": c.bytes}
        return s
