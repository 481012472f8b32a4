"""Laplace-state solves of a relaxing region, all states at once.

The reference solves a Laplace run state by state: for every s it rebuilds the
assembly at the moduli (mu(s), lam(s)) (``rebuild_for_materials``,
dense.py:650-687) and solves it (runner.py:357-391).  The operators of a
region depend on its moduli only through g = 1/(1 - nu) and mu
(docs/viscoelastic.md:68-133):

    slip -> displacement (T kernel)        T = T_A + g T_B
    traction -> displacement (U kernel)    U = (U_A + g U_B) / mu

so with u-slot columns built from T and t-slot columns from U,

    A(s) = (A0 + g(s) A1) D(s),   D(s) = 1 on u-slots, 1/mu(s) on t-slots,

and the calibrated free term (row sums of T) and a fault's right-hand side
(T times slip) are affine in g too.  ``LaplaceStates`` assembles A0 and A1
once on the device (the ordinary block kernels with the g-free and g basis
coefficients), permutes them into the block-Jacobi cluster order, and solves
every (state, right-hand side) pair as one lockstep batch
(``mbem_amd.la.batched``): y = D x solves B(g) y = b with B(g) = A0 + g A1,
then x = D^-1 y.

Scope: one region (the relaxing one) with step data carried by its faults'
slip; prescribed boundary values must be zero (``NotImplementedError``
otherwise), as in the workload this serves.  The region's closed boundary
must carry one eps (``rheology.require_one_eps``, rheology.py:527-551).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, rheology as rh
from ._lib import call, ptr, stream
from .backends.dense import AssembledDense
from .kernels import KERNEL_T, KERNEL_U
from .kernels import device as dv
from .la.batched import BatchedFGMRES, GOperator, GroupedBlockJacobi
from .la.cluster import build_cluster_tree, leaves
from .la.preconditioner import _Units
from .model.equations import rhs_terms_at

F64 = torch.float64
I64 = torch.int64

_PI = math.pi
# U = (U_A + g U_B) / mu in the basis [I1 d, eps^2 I3 d, T2[3]]:
# u_coeffs(mu, lam) = ((4 - g), 2, g) / (16 pi mu)  (basis.py:46-55)
U_SPLIT = (np.array([4.0, 2.0, 0.0]) / (16.0 * _PI), np.array([-1.0, 0.0, 1.0]) / (16.0 * _PI))
# T = T_A + g T_B in the basis [N1, N2, N3, R1, R2, R3] with N2 = N1 + 3 N3
# used to drop N2: t_coeffs(mu, lam) reduces to
# (g-1)/(4 pi) N1 + 3(g-1)/(8 pi) N3 + (4-g)/(16 pi) R1 - g/(16 pi) R2 + 3/(8 pi) R3
T_SPLIT = (np.array([-1.0 / (4 * _PI), 0.0, -3.0 / (8 * _PI), 4.0 / (16 * _PI), 0.0, 3.0 / (8 * _PI)]),
           np.array([1.0 / (4 * _PI), 0.0, 3.0 / (8 * _PI), -1.0 / (16 * _PI), -1.0 / (16 * _PI), 0.0]))


def split_coeffs(kernel: str, part: int) -> np.ndarray:
    """Basis coefficients of the g-free (part 0) or g (part 1) operator part."""
    if kernel == KERNEL_T:
        return T_SPLIT[part]
    if kernel == KERNEL_U:
        return U_SPLIT[part]
    raise NotImplementedError(f"kernel {kernel!r} has no Laplace split in mbem_amd")


@dataclass
class StatesReport:
    assembly_s: float = 0.0
    permute_s: float = 0.0
    precond_s: float = 0.0
    solve_s: float = 0.0
    rhs_s: float = 0.0
    batched: object = None
    gemm_cfg: dict = field(default_factory=dict)

    def __str__(self):
        return (f"Laplace states: assembly {self.assembly_s:.3f} s, permute {self.permute_s:.3f} s, "
                f"preconditioner {self.precond_s:.3f} s, rhs {self.rhs_s:.3f} s, "
                f"solve {self.solve_s:.3f} s; {self.batched}")


class LaplaceStates:
    """A0, A1 and the fault right-hand-side parts of one relaxing region."""

    def __init__(self, system, eps, rheology: dict, jump: str = "calibrated",
                 bj_chunk: int = 3000):
        model = system.model
        if len(model.regions) != 1:
            raise NotImplementedError("the batched Laplace route takes a single-region model")
        region = model.regions[0]
        if set(rheology) != {region.name}:
            raise ValueError(f"rheology must name the model's region {region.name!r}")
        rh.require_one_eps(model, eps, set(rheology))
        if system.rhs_values:
            raise NotImplementedError("prescribed boundary values under the Laplace route")
        for rt in system.rhs_terms:
            if rt.add_half_of_vector or rt.source_patch not in region.faults:
                raise NotImplementedError("only fault slip may load a batched Laplace run "
                                          "(prescribed boundary data must be zero)")
        _lib.require_gpu()
        self.system, self.eps, self.region = system, eps, region
        self.rheology = rh.from_spec(rheology[region.name])
        self.base = region.material
        self.layout = system.layout
        self.n = self.layout.n_unknowns
        t0 = time.perf_counter()
        self.parts = [AssembledDense(system, eps, "direct", jump=jump,
                                     _coeffs=lambda k, r, p=p: split_coeffs(k, p)) for p in (0, 1)]
        torch.cuda.synchronize()
        self.assembly_s = time.perf_counter() - t0
        # t-slot unknowns (U-kernel columns) carry D = 1/mu(s)
        tmask = np.zeros(self.n, dtype=bool)
        for s in self.layout.slots:
            if s.kind == "t":
                tmask[s.offset:s.stop] = True
        self.tmask = tmask
        # cluster order of the block-Jacobi chunks (region grouping: one super-block)
        t0 = time.perf_counter()
        units = _Units(list(self.layout.slots), self.parts[0].collocation_points_of)
        chunk_units = max(32, int(bj_chunk) // units.d)
        parts = leaves(build_cluster_tree(units.centroids, chunk_units))
        dofs = [units.dofs(np.asarray(p)) for p in parts]
        self.perm = np.concatenate(dofs).astype(np.int64)
        if np.unique(self.perm).size != self.n:
            raise RuntimeError("cluster chunks do not cover the unknowns exactly once")
        self.chunk_sizes = [d.size for d in dofs]
        dev = self.parts[0].A_dev.device
        self.device = dev
        P = torch.as_tensor(self.perm, dtype=I64, device=dev)
        self.perm_dev = P
        # A0 and A1 side by side in one n x 2L buffer (L = n rounded up to even):
        # the batched operator is then also one library GEMM [A0 | A1] [X ; X g]
        L = self.n + (self.n & 1)
        self.L = L
        self.AA = torch.zeros((self.n, 2 * L), dtype=F64, device=dev)
        self.A = []
        for k, part in enumerate(self.parts):
            view = self.AA[:, k * L:k * L + self.n]
            call("mbem_permute", ptr(part.A_dev), part.A_dev.stride(0), ptr(P), self.n, ptr(view),
                 2 * L, stream())
            part.A_dev = None
            self.A.append(view)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        self.permute_s = time.perf_counter() - t0

    # -- per-state numbers ---------------------------------------------------------
    def state_moduli(self, s):
        """(mu(s), lam(s), g(s)); s = inf is the unrelaxed region."""
        if math.isinf(s):
            mu, lam = self.base.mu, self.base.lam
        else:
            mu, lam = rh.moduli(self.rheology, self.base.mu, self.base.lam, s)
        return mu, lam, rh.g_of(mu, lam)

    def rhs_parts(self, slips) -> tuple:
        """(b0, b1), each n x R on the device: the fault right-hand side of
        slip r is b0[:, r] + g b1[:, r].  ``slips``: list of {fault name:
        (n_tri, 3) Burgers vectors}, or None entries for the model's own slip."""
        R = len(slips)
        out = [torch.zeros((self.n, R), dtype=F64, device=self.device) for _ in (0, 1)]
        part0 = self.parts[0]
        for r, slip in enumerate(slips):
            for rt in rhs_terms_at(self.system, {}):
                vec = rt.vector if slip is None else slip[rt.source_patch.name]
                x = part0._pts[id(rt.field_patch)]
                src = part0._src[id(rt.source_patch)]
                dens = torch.as_tensor(np.asarray(vec, float).reshape(-1, 3), dtype=F64,
                                       device=self.device)
                for p in (0, 1):
                    col = torch.zeros((rt.row.stop - rt.row.offset) // 3, 3, dtype=F64,
                                      device=self.device)
                    dv.t_disp(x, src, dens, split_coeffs(KERNEL_T, p), col, scale=rt.scale,
                              beta=0.0)
                    out[p][rt.row.offset:rt.row.stop, r] += col.reshape(-1)
        return out[0], out[1]

    def operator_at(self, s) -> torch.Tensor:
        """A(s) in the original unknown order (device; a check, not the solve
        path): (A0 + g A1) D(s)."""
        mu, _, g = self.state_moduli(s)
        inv = torch.empty_like(self.perm_dev)
        inv[self.perm_dev] = torch.arange(self.n, dtype=I64, device=self.device)
        Bp = self.A[0] + g * self.A[1]
        B = Bp[inv][:, inv]
        d = torch.ones(self.n, dtype=F64, device=self.device)
        d[torch.as_tensor(self.tmask, device=self.device)] = 1.0 / mu
        return B * d[None, :]

    # -- the batched solve ----------------------------------------------------------
    def solve(self, states, slips=None, rtol: float = 1e-10, restart: int = 30,
              n_anchors: int = 4, graph: bool = True, tune: bool = True, segment: int = 5):
        """Solve every (state, slip) pair.  ``states``: rheology.State tuples
        (from ``laplace_states``) or plain s values; ``slips``: list of
        {fault name: (n_tri, 3)} (None = the model's own slip).  Returns
        (x (n_states, n_slips, n) host array in the layout's unknown order,
        StatesReport)."""
        svals = [float(getattr(st, "s", st)) for st in states]
        slips = [None] if slips is None else list(slips)
        ns, nr = len(svals), len(slips)
        if ns == 0 or nr == 0:
            raise ValueError("need at least one state and one right-hand side")
        K = ns * nr
        Kp = K + (K & 1)                    # even block width (a zero column pads)
        rep = StatesReport(assembly_s=self.assembly_s, permute_s=self.permute_s)
        mods = [self.state_moduli(s) for s in svals]
        gs = np.array([m[2] for m in mods])
        mus = np.array([m[0] for m in mods])
        # columns in increasing g: state order[i] owns columns i*nr .. (i+1)*nr
        order = np.argsort(gs, kind="stable")
        gcol = np.full(Kp, gs[order[-1]])
        gcol[:K] = np.repeat(gs[order], nr)
        # anchors: na groups of consecutive states (in g), the median g of each;
        # group boundaries moved down to even columns for 16-byte loads
        na = max(1, min(int(n_anchors), ns))
        bounds = [int(round(q * ns / na)) for q in range(na + 1)]
        anchor_vals, col_ranges = [], []
        for q in range(na):
            i0, i1 = bounds[q], bounds[q + 1]
            if i1 <= i0:
                continue
            c0 = i0 * nr - ((i0 * nr) & 1)
            c1 = Kp if q == na - 1 else i1 * nr - ((i1 * nr) & 1)
            if c1 <= c0:
                continue
            anchor_vals.append(float(np.median(gs[order[i0:i1]])))
            col_ranges.append((c0, c1))
        t0 = time.perf_counter()
        M = GroupedBlockJacobi(self.A[0], self.A[1], self.chunk_sizes, anchor_vals, col_ranges, Kp)
        rep.precond_s = time.perf_counter() - t0
        # right-hand sides, permuted into cluster order
        t0 = time.perf_counter()
        b0, b1 = self.rhs_parts(slips)
        Bmat = torch.zeros((self.n, Kp), dtype=F64, device=self.device)
        for i, si in enumerate(order):
            Bmat[:, i * nr:(i + 1) * nr] = b0 + gs[si] * b1
        Bp = Bmat[self.perm_dev].contiguous()
        rep.rhs_s = time.perf_counter() - t0
        g_dev = torch.as_tensor(gcol, dtype=F64, device=self.device)
        op = GOperator(self.A[0], self.A[1], g_dev, pair=(self.AA, self.L))
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            solver = BatchedFGMRES(op, M, restart=restart, segment=segment, graph=graph)
            if tune:
                op.tune(solver.Z[0], solver.V[1])
                M.tune(solver.V[0], solver.Z[0])
            rep.gemm_cfg = {"gemm_g": op.cfg, "grouped_bj": M.cfg}
            t0 = time.perf_counter()
            brep = solver.solve(Bp, rtol=rtol)
            side.synchronize()
            rep.solve_s = time.perf_counter() - t0
            Y = solver.X[:, :K].clone()
            solver.release()
        torch.cuda.current_stream().wait_stream(side)
        rep.batched = brep
        # back to the unknown order, then x = D^-1 y
        Yo = torch.empty_like(Y)
        Yo[self.perm_dev] = Y
        xs = Yo.T.reshape(ns, nr, self.n).cpu().numpy()
        x = np.empty_like(xs)
        x[order] = xs
        t = self.tmask
        for i in range(ns):
            x[i][:, t] *= mus[i]
        self.last_report = rep
        return x, rep

    def split(self, x: np.ndarray) -> dict:
        """{slot name: (n_dofs, 3)} of one solution vector."""
        return {s.name: x[s.offset:s.stop].reshape(-1, 3) for s in self.layout.slots}

    def release(self):
        self.A = []
        self.AA = None
        for p in self.parts:
            p.release()
