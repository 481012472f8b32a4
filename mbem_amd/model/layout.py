"""Deterministic unknown layout (mbem's src/mbem/model/layout.py).

Slots: regions in model order, patches in declaration order, each patch once;
FREE_TRACTION emits its u-slot, PRESCRIBED_DISPLACEMENT its t-slot, INTERFACE
u then t (layout.py:1-24, :49-82).  A slot holds 3 * n_dofs values.  Rows:
the BIE of region R at patch q lands on q's u-slot (FREE_TRACTION), t-slot
(PRESCRIBED_DISPLACEMENT), or for an INTERFACE the u-slot for the first
incident region and the t-slot for the second (:93-108).
"""
from __future__ import annotations

from dataclasses import dataclass

from .core import BCType, Patch, Region, RegionModel


@dataclass(frozen=True)
class Slot:
    patch: Patch
    kind: str          # "u" | "t"
    offset: int
    size: int          # 3 * patch.n_dofs

    @property
    def name(self) -> str:
        return f"{self.kind}:{self.patch.name}"

    @property
    def stop(self) -> int:
        return self.offset + self.size


class UnknownLayout:
    def __init__(self, model: RegionModel):
        self.model = model
        self.slots: list = []
        self._by_key: dict = {}
        offset = 0
        emitted: set = set()
        for region in model.regions:
            for patch in region.patches:
                if id(patch) in emitted:
                    continue
                emitted.add(id(patch))
                kinds = {BCType.FREE_TRACTION: ("u",),
                         BCType.PRESCRIBED_DISPLACEMENT: ("t",),
                         BCType.INTERFACE: ("u", "t")}[patch.bc]
                for kind in kinds:
                    slot = Slot(patch, kind, offset, 3 * patch.n_dofs)
                    self.slots.append(slot)
                    self._by_key[(id(patch), kind)] = slot
                    offset += slot.size
        self.n_unknowns = offset

    def has_kind(self, kind: str) -> bool:
        return any(s.kind == kind for s in self.slots)

    def slot(self, patch: Patch, kind: str) -> Slot:
        return self._by_key[(id(patch), kind)]

    def has_slot(self, patch: Patch, kind: str) -> bool:
        return (id(patch), kind) in self._by_key

    def row_slot(self, region: Region, patch: Patch) -> Slot:
        if patch.bc is BCType.FREE_TRACTION:
            return self.slot(patch, "u")
        if patch.bc is BCType.PRESCRIBED_DISPLACEMENT:
            return self.slot(patch, "t")
        if patch.bc is BCType.INTERFACE:
            first, second = self.model.interface_regions(patch)
            if region is first:
                return self.slot(patch, "u")
            if region is second:
                return self.slot(patch, "t")
            raise ValueError(f"region '{region.name}' is not incident to interface "
                             f"'{patch.name}'")
        raise ValueError(f"no collocation row for bc {patch.bc}")
