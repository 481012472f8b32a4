"""Block-equation generation (mbem's src/mbem/model/equations.py, P0).

The sign rule (equations.py:3-20): for region R, collocation patch q and
sigma(R, p) = +1 iff p's stored normals point out of R,

    A[row(R,q), u_p] += sigma(R,p) H_qp + D_q delta_qp
    A[row(R,q), t_p] += -sigma(R,p) G_qp
    b[row(R,q)]      -= sum_f sigma(R,f) H_qf slip_f

with H the T kernel, G the U kernel, prescribed columns moved to the right-
hand side with their left-hand coefficients, and D_q the collocation free
term: COLLOCATION_JUMP * I (jump="half") or the calibrated C_q (jump=
"calibrated").  The fault-convention self-check the reference runs here
(selfcheck.ensure_fault_convention) runs in mbem_amd where the kernels
run: the GPU dense backend and the evaluators call it.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..kernels import KERNEL_T, KERNEL_U
from .core import FAULT_KINDS, BCType, Patch, Region, RegionModel
from .layout import Slot, UnknownLayout

COLLOCATION_JUMP = 0.5   # equations.py:134


@dataclass(frozen=True)
class BlockTerm:
    """A[row, col] += scale * K (+ the free term on the q == p u-block)
    (equations.py:137-164)."""
    row: Slot
    col: Slot
    kernel: str
    field_patch: Patch
    source_patch: Patch
    region: Region
    scale: float
    diag: float = 0.0
    col_op: object = None

    def col_vector(self, x: np.ndarray) -> np.ndarray:
        return x if self.col_op is None else diagonal_matvec(self.col_op, x)

    def col_block(self, blk: np.ndarray) -> np.ndarray:
        return blk if self.col_op is None else block_times_diagonal(blk, self.col_op)


@dataclass(frozen=True)
class RhsTerm:
    """b[row] += scale * K @ vector (equations.py:167-179)."""
    row: Slot
    kernel: str
    field_patch: Patch
    source_patch: Patch
    region: Region
    scale: float
    vector: np.ndarray
    add_half_of_vector: bool = False
    scale_half: float = 0.0


@dataclass(frozen=True)
class RhsValue:
    """b[row] += vector (equations.py:182-188)."""
    row: Slot
    vector: np.ndarray


@dataclass
class BlockSystem:
    model: RegionModel
    layout: UnknownLayout
    terms: list = field(default_factory=list)
    rhs_terms: list = field(default_factory=list)
    rhs_values: list = field(default_factory=list)

    def mesh_pairs(self) -> set:
        return {(id(t.field_patch), id(t.source_patch), t.kernel) for t in self.terms}

    def traction_faults(self) -> list:
        return []


def generate_system(model: RegionModel) -> BlockSystem:
    """Terms of every region's BIE (equations.py:212-309, P0 without gravity,
    slipping interfaces or traction faults)."""
    layout = UnknownLayout(model)
    system = BlockSystem(model=model, layout=layout)
    for region in model.regions:
        for f in region.faults:
            if f.bc is not BCType.FAULT:
                raise NotImplementedError(f"fault '{f.name}': only prescribed-slip FAULTs are "
                                          f"in mbem_amd")
        for q in region.patches:
            row = layout.row_slot(region, q)
            for p in region.patches:
                sigma = float(model.orientation(region, p))
                if layout.has_slot(p, "u"):
                    system.terms.append(BlockTerm(
                        row=row, col=layout.slot(p, "u"), kernel=KERNEL_T, field_patch=q,
                        source_patch=p, region=region, scale=sigma,
                        diag=COLLOCATION_JUMP if p is q else 0.0))
                else:
                    u_bar = model.value_of(p).ravel()
                    if np.any(u_bar):
                        system.rhs_terms.append(RhsTerm(
                            row=row, kernel=KERNEL_T, field_patch=q, source_patch=p,
                            region=region, scale=-sigma, vector=u_bar,
                            add_half_of_vector=(p is q), scale_half=-COLLOCATION_JUMP))
                if layout.has_slot(p, "t"):
                    system.terms.append(BlockTerm(
                        row=row, col=layout.slot(p, "t"), kernel=KERNEL_U, field_patch=q,
                        source_patch=p, region=region, scale=-sigma))
                else:
                    t_bar = model.value_of(p).ravel()
                    if np.any(t_bar):
                        system.rhs_terms.append(RhsTerm(
                            row=row, kernel=KERNEL_U, field_patch=q, source_patch=p,
                            region=region, scale=sigma, vector=t_bar))
            for f in region.faults:
                sigma = float(model.orientation(region, f))
                slip = f.value_array().ravel()
                if np.any(slip):
                    system.rhs_terms.append(RhsTerm(
                        row=row, kernel=KERNEL_T, field_patch=q, source_patch=f,
                        region=region, scale=-sigma, vector=slip))
    return system


def rhs_terms_at(system: BlockSystem, materials: dict) -> list:
    """A state's RHS terms (equations.py:405-418): with no far field they are
    the system's own."""
    return system.rhs_terms


def calibrated_diagonal(system: BlockSystem, rowsum) -> dict:
    """{(id(region), id(q)): C (n_nodes, 3, 3)} with
    C_c = -sum_p sigma(R,p) rowsum(H_qp)[c] (equations.py:512-537)."""
    model = system.model
    calib: dict = {}
    for region in model.regions:
        for q in region.patches:
            C = np.zeros((q.n_nodes, 3, 3))
            for p in region.patches:
                C -= float(model.orientation(region, p)) * rowsum(region, q, p)
            calib[(id(region), id(q))] = C
    return calib


def term_diagonal(term: BlockTerm, calib):
    """The (n_nodes, 3, 3) free term this term adds, or None
    (equations.py:540-549)."""
    if not term.diag:
        return None
    if calib is not None:
        return calib[(id(term.region), id(term.field_patch))]
    return np.broadcast_to(term.diag * np.eye(3), (term.field_patch.n_nodes, 3, 3))


def diagonal_matvec(D: np.ndarray, x: np.ndarray) -> np.ndarray:
    return np.einsum("nij,nj->ni", D, x.reshape(-1, 3)).ravel()


def block_times_diagonal(blk: np.ndarray, D: np.ndarray) -> np.ndarray:
    rows = blk.shape[0]
    return np.einsum("rnk,nkj->rnj", blk.reshape(rows, -1, 3), D).reshape(rows, -1)


def add_block_diagonal(A: np.ndarray, r0: int, c0: int, D: np.ndarray, shape=None):
    """A[r0+3c+a, c0+3c+b] += D[c,a,b] (equations.py:566-582 at P0)."""
    c = np.arange(D.shape[0])
    for a in range(3):
        for b in range(3):
            A[r0 + 3 * c + a, c0 + 3 * c + b] += D[:, a, b]


def add_jump_rhs(system: BlockSystem, calib, b: np.ndarray, rhs_terms=None):
    """Prescribed-displacement rows: the free term times the known u_bar moves
    to the RHS (equations.py:585-603)."""
    for rt in system.rhs_terms if rhs_terms is None else rhs_terms:
        if not rt.add_half_of_vector:
            continue
        r0, r1 = rt.row.offset, rt.row.stop
        u_c = rt.field_patch.collocation_values(rt.vector).ravel()
        if calib is None:
            b[r0:r1] += rt.scale_half * u_c
        else:
            b[r0:r1] -= diagonal_matvec(calib[(id(rt.region), id(rt.field_patch))], u_c)


def region_faults(model: RegionModel):
    """(region, fault) pairs in model order."""
    return [(r, f) for r in model.regions for f in r.faults if f.bc in FAULT_KINDS]
