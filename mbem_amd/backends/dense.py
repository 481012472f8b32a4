"""Dense backend on one MI355X: the system matrix assembled in HBM by the HIP
pair kernels, solved by FGMRES with the block Gauss-Seidel / block-Jacobi
preconditioner, or by LU (mbem's src/mbem/backends/dense.py).

API as the reference's: ``DenseBackend(mode, jump, deflate, lu).assemble(
system, eps) -> AssembledDense`` with ``.system .layout .eps .materials .calib
.A .b .report``, ``.solve() -> {slot name: (n_dofs, 3)}``, ``.release()`` and
``.rebuild_for_materials(material_map)`` (dense.py:49-86, :358-687).  The
matrix stays on the device (``A_dev``, ``b_dev``); ``A`` / ``b`` are host
copies made on first access.  ``solve`` takes the iterative backend's
keywords (hmat.py:608-616): ``rtol, restart, maxiter, x0, precond_max_dense,
precond_bj_chunk`` for FGMRES (the default here), or ``method="lu"`` for the
reference's dense LU.

Modes (dense.py:1-30): "direct", "basis" and "legacy" give the same operator
(legacy entrywise, basis by recombining geometry-only stacks per material);
here every mode assembles through the device kernels, and
``rebuild_for_materials`` re-integrates (about 20 ms at N = 20,000), which
costs less than the reference's recombination.  The legacy mode keeps its
restrictions: a global scalar eps, and no material rebuilds.

Assembly (dense.py:480-528): every BlockTerm block is written straight into
A by one kernel launch with its sign and material applied in-loop; the
calibrated diagonal's row sums (equations.py:512-537) are fused into the T
launches, and T blocks needed only for a row sum (a prescribed-displacement
source patch) are row-summed without being stored.  A row sum is computed
only where its C_q enters A or b -- rows with a free term (a u-slot) and
prescribed rows with a nonzero datum; ``calibrate_all_rows=True`` computes
every C_q as the reference does (the skipped ones never reach A or b).
"""
from __future__ import annotations

import math
import time
import warnings
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib, defaults
from .._lib import call, ptr, stream
from ..kernels import KERNEL_T, KERNEL_U
from ..kernels import basis as kb
from ..kernels import device as dv
from ..model.core import BCType
from ..model.equations import BlockSystem, add_jump_rhs, rhs_terms_at

F64 = torch.float64


# Streams of the assembly's concurrent block launches: a few per device for
# the life of the process (default priority).
_ASM_STREAMS: dict = {}
ASM_STREAMS_MAX = 4


def _assembly_streams(k: int) -> list:
    dev = torch.cuda.current_device()
    pool = _ASM_STREAMS.setdefault(dev, [])
    while len(pool) < k:
        pool.append(torch.cuda.Stream(device=dev))
    return pool[:k]


class DenseBackend:
    """dense.py:49-86.  ``jump="calibrated"`` annihilates constant fields
    exactly; ``deflate`` borders out the rigid translations of an all-Neumann
    model; ``lu`` names the LU flavour of ``solve(method="lu")``: "double"
    (the default), or "single" / "gpu" (la/lu.py:1-36: single-precision
    factors, every solve refined in double; both on the device here)."""

    def __init__(self, mode: str = "direct", jump: str = "calibrated", deflate: bool = False,
                 lu: str | None = None, calibrate_all_rows: bool = False):
        if mode not in ("legacy", "basis", "direct"):
            raise ValueError(mode)
        if jump not in ("half", "calibrated"):
            raise ValueError(jump)
        if lu not in (None,) + LU_HOW:
            raise ValueError(f"a dense matrix is factored {LU_HOW}, not {lu!r}")
        self.mode = mode
        self.jump = jump
        self.deflate = deflate
        self.lu = lu
        self.calibrate_all_rows = calibrate_all_rows

    def assemble(self, system: BlockSystem, eps) -> "AssembledDense":
        return AssembledDense(system, eps, self.mode, jump=self.jump, deflate=self.deflate,
                              lu=self.lu, calibrate_all_rows=self.calibrate_all_rows)


def _centroid_collocated(p) -> bool:
    """Whether the patch collocates at its triangles' centroids through the
    stock Patch and TriMesh methods (kb._source_geometry then gives its source
    arrays and collocation points in one pass, bit for bit)."""
    from ..mesh import TriMesh
    from ..model.core import Patch
    return (type(p).collocation_points is Patch.collocation_points
            and type(p).nodes is Patch.nodes and type(p.mesh) is TriMesh)


def require_anchor_or_deflate(system: BlockSystem, jump: str, deflate: bool):
    """dense.py:101-121."""
    if jump == "calibrated" and not deflate and not system.model.is_anchored() \
            and system.layout.has_kind("u"):
        raise ValueError(
            "jump='calibrated' on an all-Neumann model (no PRESCRIBED_DISPLACEMENT patch): "
            "rigid translations are an EXACT null space of the calibrated operator. Pass "
            "deflate=True (bordered / projected solve) or prescribe displacement on a patch.")


def warn_half_jump_eps(system: BlockSystem, eps, jump: str):
    """dense.py:124-147."""
    if jump != "half":
        return
    worst_name, worst = None, 0.0
    for r in system.model.regions:
        for p in list(r.patches) + list(r.faults):
            try:
                ratio = float(np.max(kb.resolve_patch_eps(eps, p) / kb.element_sizes(p.mesh)))
            except ValueError:
                continue
            if ratio > worst:
                worst_name, worst = p.name, ratio
    if worst > defaults.HALF_JUMP_MAX_EPS_OVER_H:
        warnings.warn(f"jump='half' with eps/h = {worst:.2f} on patch '{worst_name}' (limit "
                      f"{defaults.HALF_JUMP_MAX_EPS_OVER_H}): this combination was measured "
                      f"NON-convergent under h-refinement; use jump='calibrated' or a smaller eps")


def warn_collocation_near_fault(system: BlockSystem, eps):
    """dense.py:150-187."""
    from ..geometry import distance_to_mesh
    limit = defaults.COLLOCATION_FAULT_CLEARANCE_EPS
    n_close, worst, names = 0, np.inf, []
    for r in system.model.regions:
        for q in r.patches:
            x = q.collocation_points()
            for f in r.faults:
                try:
                    eps_f = kb.resolve_patch_eps(eps, f)
                except ValueError:
                    continue
                # only points within reach of the fault's bounding box can be
                # close: the exact test runs on those (same result, less work)
                fv = np.asarray(f.mesh.vertices, float)
                gap = np.maximum(0.0, np.maximum(fv.min(axis=0) - x, x - fv.max(axis=0)))
                cand = np.flatnonzero(np.linalg.norm(gap, axis=1) < limit * float(eps_f.max()))
                if cand.size == 0:
                    continue
                d, idx = distance_to_mesh(x[cand], f.mesh)
                ratio = d / eps_f[idx]
                close = int(np.sum(ratio < limit))
                if close:
                    n_close += close
                    worst = min(worst, float(ratio.min()))
                    names.append(f"{q.name}/{f.name}")
    if n_close:
        warnings.warn(f"{n_close} boundary collocation point(s) lie within {limit:g} fault-eps "
                      f"of a fault element ({', '.join(names)}; min d/eps = {worst:.2f}): the "
                      f"row is not corrected")


def warn_condition(cond: float) -> bool:
    """Warn when the dense LU's 1-norm condition estimate is above
    ``defaults.COND_WARN_THRESHOLD`` (dense.py:555-562): at that size the
    solve can miss ``defaults.SOLUTION_RTOL``.  Returns whether it warned."""
    if not cond > defaults.COND_WARN_THRESHOLD:
        return False
    warnings.warn(f"dense solve: condition estimate {cond:.2e} is above "
                  f"{defaults.COND_WARN_THRESHOLD:.0e}, so the solution may miss the "
                  f"{defaults.SOLUTION_RTOL:.0e} accuracy target (thin panels, near-fluid "
                  f"material, an un-anchored calibrated jump or a discretization resonance)")
    return True


def translation_basis(layout) -> np.ndarray:
    """Orthonormal rigid-translation basis (n, 3) on the u-slots (dense.py:190-208)."""
    if not layout.has_kind("u"):
        return np.zeros((layout.n_unknowns, 0))
    Z = np.zeros((layout.n_unknowns, 3))
    for slot in layout.slots:
        if slot.kind == "u":
            for k in range(3):
                Z[slot.offset + k:slot.stop:3, k] = 1.0
    Z /= np.linalg.norm(Z, axis=0, keepdims=True)
    return Z


def material_step(old: dict, new: dict) -> float:
    """max |log(new/old)| over the moduli of the regions that changed; infinite
    when a region appears or a modulus changes sign (hmat.py:33-52)."""
    step = 0.0
    for name, mat in new.items():
        prev = old.get(name)
        if prev is None:
            return math.inf
        for a, b in ((mat.mu, prev.mu), (mat.lam, prev.lam)):
            if a == b:
                continue
            if a <= 0.0 or b <= 0.0:
                return math.inf
            step = max(step, abs(math.log(a / b)))
    return step


class _Lineage:
    """What one chain of ``rebuild_for_materials`` shares across its solves
    (hmat.py:55-70): the last solution (the warm start of
    ``x0="previous"``) and the recycle space with the materials it was last
    harvested at."""

    def __init__(self):
        self.x = None
        self.recycle = None
        self.recycle_materials = None


LU_HOW = ("double", "single", "gpu")      # la/lu.py:49


class _LUFactors:
    """Device LU factors of a row-major matrix M (mbem_lu_factor, or
    mbem_lu_factor_f32 for an FP32 copy) with what the blocked solves need:
    the inverses of the 64 x 64 diagonal blocks of L and U and the pivots as
    a row order (P^T v = v[perm])."""

    def __init__(self, LU: torch.Tensor, ipiv: torch.Tensor, m: int):
        self.LU, self.ipiv, self.m = LU, ipiv, m
        self.f32 = LU.dtype == torch.float32
        dev = LU.device
        k = int(_lib.load().mbem_lu_inv_doubles(m))
        self.inv_l = torch.empty(max(k, 1), dtype=F64, device=dev)
        self.inv_u = torch.empty(max(k, 1), dtype=F64, device=dev)
        call("mbem_lu_prepare_f32" if self.f32 else "mbem_lu_prepare", ptr(LU), m,
             ptr(self.inv_l), ptr(self.inv_u), stream())
        piv = ipiv.cpu().numpy() - 1            # rocSOLVER pivots are 1-based
        perm = np.arange(m, dtype=np.int64)
        for i, p in enumerate(piv.tolist()):
            if p != i:
                perm[i], perm[p] = perm[p], perm[i]
        self.perm = torch.as_tensor(perm, device=dev)
        self.work = torch.empty(2 * max(m, 1), dtype=F64, device=dev)

    def solve_into(self, b: torch.Tensor, x: torch.Tensor, adjoint: bool = False):
        """x = M^-1 b (adjoint: M^-T b); b and x may be the same vector."""
        call("mbem_lu_solve_blocked_f32" if self.f32 else "mbem_lu_solve_blocked", ptr(self.LU),
             self.m, ptr(self.perm), ptr(self.inv_l), ptr(self.inv_u), ptr(b), ptr(x),
             ptr(self.work), int(adjoint), stream())
        return x


def refine(solve_into, matvec_into, b: torch.Tensor, tol: float = None, passes: int = None,
           max_rate: float = None):
    """Classical refinement of a single-precision solve (la/lu.py:198-239):
    x = solve(b), then x += solve(b - M x) until what is left of the error,
    rate / (1 - rate) |d| / |x| with rate = |d| / |d_prev| (the first
    correction against the solution itself), or the correction itself, is
    at most ``tol``; at most ``passes`` corrections; stop early when a
    correction contracts slower than ``max_rate``.  Returns (x, report) with
    report = {done, passes, correction, left, rate}."""
    tol = defaults.LU_REFINE_TOL if tol is None else tol
    passes = defaults.LU_REFINE_PASSES if passes is None else passes
    max_rate = defaults.LU_REFINE_RATE if max_rate is None else max_rate
    x = torch.empty_like(b)
    solve_into(b, x)
    xn = float(torch.linalg.vector_norm(x))
    done = xn == 0.0
    r = torch.empty_like(b)
    d = torch.empty_like(b)
    before, rate, size, left, k = None, 0.0, 0.0, 0.0, 0
    while not done and k < passes:
        matvec_into(x, r)                          # r = M x
        torch.sub(b, r, out=r)                     # r = b - M x
        solve_into(r, d)
        x += d
        k += 1
        now = float(torch.linalg.vector_norm(d)) / float(torch.linalg.vector_norm(x))
        rate = now if before is None else (now / before if before > 0.0 else 0.0)
        size, before = now, now
        left = rate / (1.0 - rate) * size if rate < 1.0 else float("inf")
        done = left <= tol or size <= tol
        if rate > max_rate:
            break
    return x, {"done": bool(done), "passes": k, "correction": size, "left": left, "rate": rate}


@dataclass
class DenseSolveReport:
    """dense.py:260-271 (LU solve)."""
    cond_estimate: float
    lu: dict | None = None
    residual: float = float("nan")

    def __str__(self):
        return f"dense LU: cond ~ {self.cond_estimate:.2e}"


@dataclass
class AssemblyReport:
    """Where the assembly time went (seconds) and how many pairs ran.
    ``seconds`` is the device time from the first assembly launch to the
    last, read from two HIP events on first access (the assembly returns
    without waiting for its kernels, so the preconditioner build's launches
    can queue behind them)."""
    _events: tuple
    t_pairs_stored: int
    t_pairs_rowsum_only: int
    u_pairs: int
    rhs_pairs: int

    @property
    def seconds(self) -> float:
        e0, e1 = self._events
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3


class _Timings(dict):
    """Phase timings; ``assembly_s`` is read from the assembly's events when
    first looked at."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self._resolve = None

    def _fill(self):
        if self._resolve is not None and dict.get(self, "assembly_s") is None:
            dict.__setitem__(self, "assembly_s", self._resolve())

    def __getitem__(self, k):
        self._fill()
        return dict.__getitem__(self, k)

    def get(self, k, default=None):
        self._fill()
        return dict.get(self, k, default)

    def items(self):
        self._fill()
        return dict.items(self)

    def values(self):
        self._fill()
        return dict.values(self)

    def __repr__(self):
        self._fill()
        return dict.__repr__(self)


class AssembledDense:
    def __init__(self, system: BlockSystem, eps, mode: str = "direct", jump: str = "calibrated",
                 deflate: bool = False, lu=None, calibrate_all_rows: bool = False,
                 fault_drag=None, _materials=None, _coeffs=None):
        if mode not in ("legacy", "basis", "direct"):
            raise ValueError(mode)
        if jump not in ("half", "calibrated"):
            raise ValueError(jump)
        if lu not in (None,) + LU_HOW:
            raise ValueError(f"a dense matrix is factored {LU_HOW}, not {lu!r}")
        if mode == "legacy" and isinstance(eps, (dict, str)):
            raise ValueError("legacy mode supports only global scalar eps")
        if fault_drag:
            raise NotImplementedError("fault drag needs FAULT_TRACTION faults (not in mbem_amd)")
        _lib.require_gpu()
        t0 = time.perf_counter()
        require_anchor_or_deflate(system, jump, deflate)
        self.timings = _Timings(checks_s=time.perf_counter() - t0)
        self.system = system
        self.layout = system.layout
        self.eps = eps
        self.mode = mode
        self.jump = jump
        self.deflate = deflate
        self.lu = lu
        self.calibrate_all_rows = calibrate_all_rows
        # (kernel, region) -> basis coefficients; None = from the region's
        # moduli.  The Laplace-state split assembles the g-free and g parts
        # of the operator through it.
        self._coeff_fn = _coeffs
        self.materials = dict(_materials) if _materials is not None else \
            {r.name: r.material for r in system.model.regions}
        self.fault_drag = {}
        self.report = None
        self._precond = None
        self._precond_caps = None
        self._lu = None
        self._lineage = _Lineage()
        self._A_host = None
        self._b_host = None
        self.A_dev = None
        self.b_dev = None
        t0 = time.perf_counter()
        self._prepare_geometry()
        self.timings["geometry_s"] = time.perf_counter() - t0
        self.timings["assembly_s"] = None
        self._build()
        # the two checks that only warn run while the device assembles (the
        # build returns once its kernels are queued)
        t0 = time.perf_counter()
        warn_half_jump_eps(system, eps, jump)
        warn_collocation_near_fault(system, eps)
        self.timings["checks_s"] += time.perf_counter() - t0
        # the resolver holds the report, not ``self``: no reference cycle keeps
        # a dropped assembly (and its n x n matrix) alive until a full GC
        rep = self.assembly
        self.timings._resolve = lambda: rep.seconds

    # -- geometry on the device ------------------------------------------
    def _prepare_geometry(self):
        from ..selfcheck import ensure_fault_convention
        if any(r.faults for r in self.system.model.regions):
            ensure_fault_convention()
        dev = _lib.device()
        self.device = dev
        self._src: dict = {}
        self._pts: dict = {}
        self._pts_host: dict = {}
        self._eps_arr: dict = {}
        patches = []
        for r in self.system.model.regions:
            for p in list(r.patches) + list(r.faults):
                if id(p) not in self._src:
                    patches.append(p)
                    eps_arr = kb.resolve_patch_eps(self.eps, p)
                    if _centroid_collocated(p):
                        tv, nrm, x = kb._source_geometry(p.mesh)
                    else:
                        tv, nrm = kb._source_arrays(p.mesh)
                        x = np.ascontiguousarray(p.collocation_points())
                    self._eps_arr[id(p)] = eps_arr
                    # pinned, non-blocking uploads (the staging ring)
                    self._src[id(p)] = dv.Sources(_lib.upload(tv, torch.float64),
                                                  _lib.upload(nrm, torch.float64),
                                                  _lib.upload(eps_arr, torch.float64))
                    self._pts_host[id(p)] = x
                    self._pts[id(p)] = _lib.upload(x, torch.float64)
        self._patches = patches

    def collocation_points_of(self, patch) -> np.ndarray:
        return self._pts_host[id(patch)]

    def eps_for(self, patch) -> np.ndarray:
        return self._eps_arr[id(patch)]

    # -- assembly -----------------------------------------------------------
    def _coeffs(self, kernel, region):
        if self._coeff_fn is not None:
            return np.asarray(self._coeff_fn(kernel, region), dtype=float)
        m = self.materials[region.name]
        return kb.t_coeffs(m.mu, m.lam) if kernel == KERNEL_T else kb.u_coeffs(m.mu, m.lam)

    def _calib_needed(self):
        """(region, q) pairs whose calibrated C_q enters A or b."""
        need = set()
        for t in self.system.terms:
            if t.diag:
                need.add((id(t.region), id(t.field_patch)))
        for rt in self.system.rhs_terms:
            if rt.add_half_of_vector:
                need.add((id(rt.region), id(rt.field_patch)))
        if self.calibrate_all_rows:
            for r in self.system.model.regions:
                for q in r.patches:
                    need.add((id(r), id(q)))
        return need

    def _build(self):
        torch.cuda.synchronize()
        n = self.layout.n_unknowns
        dev = self.device
        model = self.system.model
        # every host input goes to the device before the first kernel, so no
        # copy below waits on the assembly and the host returns at once
        rhs_in = []
        for rt in rhs_terms_at(self.system, self.materials):
            rhs_in.append((rt, torch.as_tensor(np.asarray(rt.vector, float).reshape(-1, 3),
                                               dtype=F64, device=dev)))
        rv_in = [(rv, torch.as_tensor(rv.vector, dtype=F64, device=dev))
                 for rv in self.system.rhs_values]
        jump_in = []
        for rt in self.system.rhs_terms:
            if rt.add_half_of_vector:
                u_c = np.ascontiguousarray(rt.field_patch.collocation_values(rt.vector).ravel())
                nq = u_c.size // 3
                jump_in.append((rt, torch.as_tensor(u_c, dtype=F64, device=dev),
                                torch.arange(nq + 1, dtype=torch.int64, device=dev)))
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        A = torch.empty((n, n), dtype=F64, device=dev)
        lda = A.stride(0)
        n_ts = n_tr = n_u = n_rhs = 0
        # calibration buffers (C_q, -sum_p sigma rowsum), zero-initialised
        self.calib = None
        calib_dev = {}
        need = set()
        if self.jump == "calibrated":
            need = self._calib_needed()
            for r in model.regions:
                for q in r.patches:
                    if (id(r), id(q)) in need:
                        calib_dev[(id(r), id(q))] = torch.zeros((q.n_nodes, 3, 3), dtype=F64,
                                                                device=dev)
        b = torch.zeros(n, dtype=F64, device=dev)
        # Concurrent launches: each block kernel is only a round or two of
        # waves deep, so on one stream every block ends in a tail of idle
        # SIMDs.  Blocks of different field patches (collocation rows) write
        # disjoint rows of A and b and disjoint row-sum buffers, so each field
        # patch gets a stream of its own and the others' blocks fill its
        # tails; blocks of one field patch (shared row sums, accumulated
        # blocks, the free term after its row sums) keep their order.
        main = torch.cuda.current_stream()
        rows = []
        for item in list(self.system.terms) + [rt for rt, _ in rhs_in]:
            if id(item.field_patch) not in [id(p) for p in rows]:
                rows.append(item.field_patch)
        for r in model.regions:
            for q in r.patches:
                if (id(r), id(q)) in calib_dev and id(q) not in [id(p) for p in rows]:
                    rows.append(q)
        pool = _assembly_streams(max(1, min(ASM_STREAMS_MAX, len(rows))))
        row_stream = {id(p): pool[i % len(pool)] for i, p in enumerate(rows)}
        ready = torch.cuda.Event()
        ready.record(main)
        for s_ in pool:
            s_.wait_event(ready)

        def on(patch):
            return torch.cuda.stream(row_stream[id(patch)])

        covered = set()
        stored_T = set()
        for t in self.system.terms:
            r0, c0 = t.row.offset, t.col.offset
            key = (t.row.name, t.col.name)
            beta = 1.0 if key in covered else 0.0
            covered.add(key)
            x = self._pts[id(t.field_patch)]
            src = self._src[id(t.source_patch)]
            out = A[r0:, c0:]
            c = self._coeffs(t.kernel, t.region)
            with on(t.field_patch):
                if t.kernel == KERNEL_T:
                    ck = (id(t.region), id(t.field_patch))
                    rs = calib_dev.get(ck) if (ck in calib_dev) else None
                    dv.t_block(x, src, c, out, lda, scale=t.scale, beta=beta, rowsum=rs,
                               rs_scale=-float(model.orientation(t.region, t.source_patch)))
                    stored_T.add((id(t.region), id(t.field_patch), id(t.source_patch)))
                    n_ts += x.shape[0] * src.ns
                elif t.kernel == KERNEL_U:
                    dv.u_block(x, src, c, out, lda, scale=t.scale, beta=beta)
                    n_u += x.shape[0] * src.ns
                else:
                    raise NotImplementedError(f"kernel {t.kernel!r} is not in mbem_amd")
        # zero the blocks no term writes
        slots = self.layout.slots
        for rs_ in slots:
            for cs_ in slots:
                if (rs_.name, cs_.name) not in covered:
                    A[rs_.offset:rs_.stop, cs_.offset:cs_.stop].zero_()
        # row sums of T blocks that are not stored (prescribed-u source patches)
        if self.jump == "calibrated":
            for r in model.regions:
                for q in r.patches:
                    ck = (id(r), id(q))
                    if ck not in calib_dev:
                        continue
                    for p in r.patches:
                        if (id(r), id(q), id(p)) in stored_T:
                            continue
                        x = self._pts[id(q)]
                        src = self._src[id(p)]
                        with on(q):
                            dv.t_block(x, src, self._coeffs(KERNEL_T, r), None, 0,
                                       rowsum=calib_dev[ck],
                                       rs_scale=-float(model.orientation(r, p)), store=False)
                        n_tr += x.shape[0] * src.ns
            self.calib_dev = calib_dev
        # free terms on the u-block diagonals
        for t in self.system.terms:
            if not t.diag:
                continue
            blk = A[t.row.offset:, t.col.offset:]
            nq = t.field_patch.n_nodes
            with on(t.field_patch):
                if self.jump == "calibrated":
                    D = calib_dev[(id(t.region), id(t.field_patch))]
                else:
                    D = torch.as_tensor(np.broadcast_to(t.diag * np.eye(3), (nq, 3, 3)).copy(),
                                        dtype=F64, device=dev)
                dv.add_block_diag(blk, lda, D)
        # right-hand side (dense.py:514-524)
        for rt, dens in rhs_in:
            x = self._pts[id(rt.field_patch)]
            src = self._src[id(rt.source_patch)]
            out = b[rt.row.offset:rt.row.stop].view(-1, 3)
            c = self._coeffs(rt.kernel, rt.region)
            with on(rt.field_patch):
                if rt.kernel == KERNEL_T:
                    dv.t_disp(x, src, dens, c, out, scale=rt.scale, beta=1.0)
                else:
                    dv.u_disp(x, src, dens, c, out, scale=rt.scale, beta=1.0)
            n_rhs += x.shape[0] * src.ns
        # join: the rest runs on the caller's stream behind every block
        for s_ in pool:
            done = torch.cuda.Event()
            done.record(s_)
            main.wait_event(done)
        for rv, vec in rv_in:
            b[rv.row.offset:rv.row.stop] += vec
        # prescribed-displacement rows: the free term times u_bar moves to the
        # right-hand side (equations.py:585-603), b -= blockdiag(C_q) u_bar on
        # the device (a block-CSR product with one diagonal block per row)
        for rt, u_c, ident in jump_in:
            r0, r1 = rt.row.offset, rt.row.stop
            if self.jump != "calibrated":
                b[r0:r1] += rt.scale_half * u_c
            else:
                C = calib_dev[(id(rt.region), id(rt.field_patch))]
                call("mbem_bsr_gemv", (r1 - r0) // 3, ptr(ident), ptr(ident), ptr(C), ptr(u_c),
                     ptr(b[r0:r1]), -1.0, stream())
        ev1.record()
        self.A_dev = A
        self.b_dev = b
        self.assembly = AssemblyReport((ev0, ev1), n_ts, n_tr, n_u, n_rhs)
        if self.jump == "calibrated":
            self.calib = _LazyCalib(calib_dev)

    # -- host views ----------------------------------------------------------
    @property
    def A(self) -> np.ndarray:
        if self._A_host is None:
            self._A_host = self.A_dev.cpu().numpy()
        return self._A_host

    @property
    def b(self) -> np.ndarray:
        if self._b_host is None:
            self._b_host = self.b_dev.cpu().numpy()
        return self._b_host

    # -- operators -----------------------------------------------------------
    def apply_into(self, v, out):
        call("mbem_gemv", ptr(self.A_dev), self.A_dev.stride(0), self.A_dev.shape[0],
             self.A_dev.shape[1], ptr(v), ptr(out), 1.0, 0.0, stream())
        return out

    def matvec(self, v):
        out = torch.empty_like(v)
        return self.apply_into(v, out)

    # -- solve -----------------------------------------------------------------
    def solve(self, method: str = "fgmres", rtol: float = defaults.GMRES_RTOL,
              restart: int = defaults.GMRES_RESTART, maxiter: int = defaults.GMRES_MAXITER,
              x0=None, precond_max_dense=None,
              precond_above_dense: str = defaults.PRECOND_RUNG_ABOVE_DENSE,
              precond_hodlr_max=None, precond_bj_chunk=None, recycle=None,
              return_device: bool = False, precond_grouping: str = defaults.PRECOND_GROUPING,
              precond_coarse: str = None):
        """{slot name: (n_dofs, 3)}; leaves ``self.report`` (SolveReport for
        FGMRES, DenseSolveReport for LU).  ``precond_grouping`` ("patch": a
        Gauss-Seidel sweep over patches, the reference's default; "region":
        one super-block per region, i.e. plain block-Jacobi over the cluster
        chunks) is an addition to the reference's keywords.

        The FGMRES route takes the compressed backend's two sequence options
        for the solves of one ``rebuild_for_materials`` chain
        (hmat.py:605-664): ``x0="previous"`` warm-starts from the chain's last
        solution, and ``recycle=True`` (default
        ``defaults.GCRO_RECYCLE_DEFAULT``) carries a GCRO-DR subspace through
        it, dropped when the materials have moved more than
        ``defaults.PRECOND_REUSE_MAX_STEP`` since it was harvested.  A
        ``RecycleSpace`` may also be passed to carry one explicitly.

        ``precond_coarse`` ("const", "rigid" or "linear"; an addition to
        the reference's keywords) puts the coarse correction of
        la/coarse.py over the block preconditioner."""
        lineage = self._lineage
        if isinstance(x0, str):
            if x0 != "previous":
                raise ValueError(f"x0 must be an array, None or 'previous' (got {x0!r})")
            x0 = lineage.x if (lineage.x is not None
                               and lineage.x.shape[0] == self.layout.n_unknowns) else None
        rs = None
        if method == "fgmres":
            from ..la.solver import RecycleSpace
            if isinstance(recycle, RecycleSpace):
                rs = recycle
            elif defaults.GCRO_RECYCLE_DEFAULT if recycle is None else recycle:
                if lineage.recycle is None:
                    lineage.recycle = RecycleSpace()
                elif lineage.recycle_materials is not None and material_step(
                        lineage.recycle_materials, self.materials) > \
                        defaults.PRECOND_REUSE_MAX_STEP:
                    lineage.recycle.reset()
                rs = lineage.recycle
                lineage.recycle_materials = dict(self.materials)
        recycle = rs
        if method == "lu":
            x = self._solve_lu()
        elif method == "fgmres":
            x = self._solve_fgmres(rtol, restart, maxiter, x0, precond_max_dense,
                                   precond_above_dense, precond_bj_chunk, recycle,
                                   precond_grouping, precond_coarse)
        else:
            raise ValueError(f"solve method {method!r} (expected 'fgmres' or 'lu')")
        self.x_dev = x
        lineage.x = x.detach().clone()
        if return_device:
            return x
        xh = x.cpu().numpy()
        return {s.name: xh[s.offset:s.stop].reshape(-1, 3) for s in self.layout.slots}

    def preconditioner(self, max_dense=None, bj_chunk=None,
                       above_dense=defaults.PRECOND_RUNG_ABOVE_DENSE,
                       grouping=defaults.PRECOND_GROUPING, coarse=None):
        from ..la.preconditioner import BlockGaussSeidel
        caps = (max_dense, above_dense, bj_chunk, grouping, coarse)
        if self._precond is not None and self._precond_caps != caps:
            self._precond = None
        if self._precond is None:
            M = BlockGaussSeidel(self, max_dense=max_dense, bj_chunk=bj_chunk,
                                 above_dense=above_dense, grouping=grouping)
            if coarse:
                from ..la.coarse import CoarseSpace, TwoLevel
                M = TwoLevel(M, CoarseSpace(self, M.chunk_dofs, basis=coarse))
            self._precond = M
            self._precond_caps = caps
        return self._precond

    def _solve_fgmres(self, rtol, restart, maxiter, x0, max_dense, above_dense, bj_chunk,
                      recycle, grouping=defaults.PRECOND_GROUPING, coarse=None):
        from ..la.solver import fgmres
        t0 = time.perf_counter()
        M = self.preconditioner(max_dense, bj_chunk, above_dense, grouping, coarse)
        self.timings["precond_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        if x0 is not None and not isinstance(x0, torch.Tensor):
            x0 = torch.as_tensor(np.asarray(x0, float), dtype=F64, device=self.device)
        if self.deflate:
            Z = torch.as_tensor(translation_basis(self.layout), dtype=F64, device=self.device)

            def proj(v):
                return v - Z @ (Z.T @ v)

            class _P:
                def __init__(s, f):
                    s.f = f

                def apply_into(s, v, out):
                    out.copy_(proj(s.f(proj(v))))
                    return out

            x, report = fgmres(_P(self.matvec), proj(self.b_dev), M=_P(M), x0=x0, rtol=rtol,
                               restart=restart, maxiter=maxiter, recycle=recycle)
            x = proj(x)
        else:
            x, report = fgmres(self, self.b_dev, M=M, x0=x0, rtol=rtol, restart=restart,
                               maxiter=maxiter, recycle=recycle)
        torch.cuda.synchronize()
        self.timings["fgmres_s"] = time.perf_counter() - t0
        report.precond_summary = M.summary()
        self.report = report
        return x

    def _solve_lu(self):
        """LU with partial pivoting on the device (dense.py:532-571, lu.py:110-161)."""
        n = self.A_dev.shape[0]
        dev = self.device
        if self._lu is None:
            Z = translation_basis(self.layout) if self.deflate else None
            if Z is not None and Z.shape[1] == 0:
                Z = None
            if Z is not None:
                K = torch.zeros((n + 3, n + 3), dtype=F64, device=dev)
                K[:n, :n] = self.A_dev
                Zd = torch.as_tensor(Z, dtype=F64, device=dev)
                K[:n, n:] = Zd
                K[n:, :n] = Zd.T
                LU = K
            else:
                LU = self.A_dev
            m = LU.shape[0]
            how = self.lu or defaults.DENSE_LU
            t0 = time.perf_counter()
            anorm = self._norm1(LU)
            if how == "double":
                if LU is self.A_dev:
                    LU = LU.clone()                # getrf overwrites; A stays for the residual
                fac = self._factor(LU, "mbem_lu_factor")
            else:
                # single-precision factors of an FP32 copy; the FP64 matrix is
                # kept (only read) for the refinement's residuals
                fac = self._factor(LU.to(torch.float32), "mbem_lu_factor_f32")
            fs = time.perf_counter() - t0
            self._lu = (fac, fs, self._cond1(fac, anorm), {"how": how, "factor_s": fs,
                                                           "M": LU})
            warn_condition(self._lu[2])
        from ..la.solver import gemv
        fac, fs, cond, lu_rep = self._lu
        m = fac.m
        rhs = torch.zeros(m, dtype=F64, device=dev)
        rhs[:n] = self.b_dev
        if not fac.f32:
            fac.solve_into(rhs, rhs)
            y = rhs
        else:
            M = lu_rep["M"]
            y, ref = refine(lambda v, out: fac.solve_into(v, out),
                            lambda v, out: gemv(M, v, out), rhs)
            lu_rep["passes"] = max(ref["passes"], lu_rep.get("passes", 0))
            if ref["left"] >= lu_rep.get("left", 0.0):
                lu_rep["correction"], lu_rep["left"] = ref["correction"], ref["left"]
            if not ref["done"]:
                warnings.warn(f"dense LU in single precision ({lu_rep['how']}): the refinement "
                              f"did not contract (rate {ref['rate']:.1e}, {ref['left']:.1e} "
                              f"left after {ref['passes']} corrections); factoring in double. "
                              f"The matrix is too ill-conditioned for single-precision factors")
                t0 = time.perf_counter()
                fac = self._factor(M.clone(), "mbem_lu_factor")
                lu_rep.update(fallback=True, factor_s=lu_rep["factor_s"] + time.perf_counter() - t0)
                self._lu = (fac, fs, cond, lu_rep)
                fac.solve_into(rhs, rhs)
                y = rhs
        x = y[:n].contiguous()
        r = self.b_dev - self.matvec(x)
        res = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(self.b_dev))
        self.report = DenseSolveReport(cond_estimate=cond,
                                       lu={k: v for k, v in lu_rep.items() if k != "M"},
                                       residual=res)
        return x

    @staticmethod
    def _factor(M: torch.Tensor, entry: str) -> "_LUFactors":
        """getrf of M in place (FP64 or FP32 by ``entry``) and the blocked-solve
        tables; raises on a singular factor."""
        m = M.shape[0]
        ipiv = torch.empty(m, dtype=torch.int64, device=M.device)
        info = torch.zeros(1, dtype=torch.int64, device=M.device)
        call(entry, ptr(M), m, ptr(ipiv), ptr(info), stream())
        if int(info.item()) != 0:
            raise np.linalg.LinAlgError(f"dense LU: singular (getrf info {int(info.item())})")
        return _LUFactors(M, ipiv, m)

    @staticmethod
    def _norm1(M: torch.Tensor) -> float:
        """max_c sum_r |M[r, c]| of a square device matrix (la/lu.py:142-145)."""
        m = M.shape[0]
        colsum = torch.empty(m, dtype=F64, device=M.device)
        work = torch.empty(int(_lib.load().mbem_colsum_work(m)), dtype=F64, device=M.device)
        call("mbem_colsum_abs", ptr(M), M.stride(0), m, m, ptr(work), ptr(colsum), stream())
        return float(colsum.max().item()) if m else 0.0

    @staticmethod
    def _cond1(fac: "_LUFactors", anorm: float) -> float:
        """1-norm condition estimate of the factored matrix: gecon's
        rcond = (1 / est ||M^-1||_1) / ||M||_1 from the device solves, then
        1 / rcond, infinite when rcond is 0 (la/lu.py:241-248)."""
        from ..la.condest import onenorm_inverse
        v = torch.empty(fac.m, dtype=F64, device=fac.LU.device)

        def solve(x, adjoint):
            v.copy_(torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)))
            fac.solve_into(v, v, adjoint=adjoint)
            return v.cpu().numpy()

        ainv = onenorm_inverse(lambda x: solve(x, False), lambda x: solve(x, True), fac.m)
        if ainv == 0.0 or anorm == 0.0:
            return float("inf")
        rcond = (1.0 / ainv) / anorm
        return (1.0 / rcond) if (np.isfinite(rcond) and rcond > 0) else float("inf")

    def release(self):
        """Drop the factors and the preconditioner (dense.py:573-576)."""
        self._lu = None
        self._precond = None

    def rebuild_for_materials(self, material_map: dict, fault_drag=None) -> "AssembledDense":
        """A new assembly of the same system with some regions' materials
        replaced (dense.py:650-687), re-integrated on the device."""
        if self.mode == "legacy":
            raise ValueError("rebuild_for_materials requires mode='basis' or 'direct'")
        mats = dict(self.materials)
        for name, mat in material_map.items():
            if name not in mats:
                raise KeyError(f"unknown region '{name}'")
            mats[name] = mat
        new = AssembledDense(self.system, self.eps, self.mode, jump=self.jump,
                             deflate=self.deflate, lu=self.lu,
                             calibrate_all_rows=self.calibrate_all_rows, _materials=mats)
        new._lineage = self._lineage            # one chain, one layout
        return new


class _LazyCalib(dict):
    """``calib`` as the reference exposes it ({(id(region), id(q)): (n,3,3)}),
    copied to the host on first access."""

    def __init__(self, dev_map):
        super().__init__()
        self._dev = dev_map

    def _fill(self):
        if not super().__len__() and self._dev:
            for k, v in self._dev.items():
                super().__setitem__(k, v.cpu().numpy())

    def __getitem__(self, k):
        self._fill()
        return super().__getitem__(k)

    def __contains__(self, k):
        return k in self._dev

    def __len__(self):
        return len(self._dev)

    def keys(self):
        self._fill()
        return super().keys()

    def items(self):
        self._fill()
        return super().items()

    def values(self):
        self._fill()
        return super().values()
