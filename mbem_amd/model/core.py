"""Region graph with automatic boundary orientation (mbem's
src/mbem/model/core.py), for the P0 discontinuous patches of the dense hot
path.

``sigma(R, p) = +1`` iff patch p's stored normals point out of region R,
inferred from the signed solid angle the patch subtends at the region's probe
point (core.py:713-729) and validated by the Gauss closure identity
``sum_p sigma(R,p) Omega_p(x_R) = 4 pi`` (core.py:785-806).  Faults answer
``FAULT_ORIENTATION = -1`` (core.py:92-100): their slip enters every site
through the same ``-sigma`` coefficient a prescribed displacement does.

Not in mbem_amd (each raises NotImplementedError where it is asked for):
Lagrange orders 1 and 2, continuous layouts, FAULT_TRACTION faults,
interfaces that slip, gravity and far-field windows.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field, replace

import numpy as np

from .. import defaults


class BCType(enum.Enum):
    """Boundary conditions (core.py:72-77)."""
    FREE_TRACTION = "free_traction"
    PRESCRIBED_DISPLACEMENT = "prescribed_displacement"
    INTERFACE = "interface"
    FAULT = "fault"
    FAULT_TRACTION = "fault_traction"


FAULT_KINDS = (BCType.FAULT, BCType.FAULT_TRACTION)
FAULT_ORIENTATION = -1
_ON_PATCH_TOL = 1e-9


@dataclass(eq=False)
class Patch:
    """A named triangulated piece with a boundary condition (core.py:103-498).

    ``value`` is the prescribed quantity (traction, displacement or the
    Burgers vector b = u(+n) - u(-n) of a FAULT): None (zeros), a (3,)
    constant, an (N_tri, 3) array, or a callable of the element centroids.
    At P0 the node, the dof node and the collocation point of an element are
    its centroid.
    """
    name: str
    mesh: object
    bc: BCType
    value: object = None
    order: int = 0
    layout: str = "discontinuous"
    allow_opening: bool = False
    locked: object = None
    slip: object = None

    def __post_init__(self):
        if self.order != int(self.order):
            raise ValueError(f"patch '{self.name}': order must be an integer 0, 1 or 2 "
                             f"(got {self.order!r})")
        self.order = int(self.order)
        if self.order not in (0, 1, 2):
            raise ValueError(f"patch '{self.name}': order must be 0, 1 or 2")
        if self.order != 0:
            raise NotImplementedError(f"patch '{self.name}': Lagrange order {self.order} is not "
                                      f"in mbem_amd (P0 only)")
        if self.layout not in ("discontinuous", "continuous"):
            raise ValueError(f"patch '{self.name}': layout must be 'discontinuous' or "
                             f"'continuous' (got {self.layout!r})")
        if self.layout == "continuous":
            raise ValueError(f"patch '{self.name}': a continuous layout is P2 only")
        if self.bc is BCType.FAULT_TRACTION:
            raise NotImplementedError(f"patch '{self.name}': FAULT_TRACTION faults (solved slip, "
                                      f"traction rows) are not in mbem_amd")
        if self.allow_opening or self.locked is not None:
            raise ValueError(f"patch '{self.name}': allow_opening and locked apply to a "
                             f"FAULT_TRACTION fault only")
        if self.slip is not None:
            if self.slip not in ("prescribed", "solved"):
                raise ValueError(f"patch '{self.name}': slip must be 'prescribed' or 'solved'")
            if self.bc is not BCType.INTERFACE:
                raise ValueError(f"patch '{self.name}': slip is of an INTERFACE patch")
            raise NotImplementedError(f"patch '{self.name}': interfaces that slip are not in "
                                      f"mbem_amd")
        self._refuse_degenerate()

    def _refuse_degenerate(self) -> None:
        """A collapsed element (height / longest edge <= MIN_TRIANGLE_HEIGHT_OVER_L)
        is refused (core.py:227-254)."""
        tv = np.asarray(self.mesh.vertices, float)[np.asarray(self.mesh.triangles)]
        L = np.linalg.norm(tv[:, [1, 2, 0], :] - tv, axis=2).max(axis=1)
        two_area = np.linalg.norm(np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0]), axis=1)
        h_over_L = two_area / np.where(L > 0.0, L, 1.0) ** 2
        bad = np.nonzero(h_over_L <= defaults.MIN_TRIANGLE_HEIGHT_OVER_L)[0]
        if bad.size:
            worst = int(bad[np.argmin(h_over_L[bad])])
            raise ValueError(
                f"patch '{self.name}': {bad.size} of {tv.shape[0]} elements are degenerate "
                f"(height / longest edge at or below {defaults.MIN_TRIANGLE_HEIGHT_OVER_L:g}; "
                f"worst {h_over_L[worst]:.1e} at element {worst})")

    @property
    def n_triangles(self) -> int:
        return self.mesh.n_triangles

    @property
    def n_nodes(self) -> int:
        return self.mesh.n_triangles

    @property
    def n_dofs(self) -> int:
        return self.mesh.n_triangles

    @property
    def continuous(self) -> bool:
        return False

    @property
    def solves_slip(self) -> bool:
        return False

    def column_ids(self):
        return None

    def expand(self, values) -> np.ndarray:
        v = np.asarray(values, float)
        if v.shape != (self.n_nodes, 3):
            raise ValueError(f"patch '{self.name}': values must be ({self.n_nodes}, 3), "
                             f"got {v.shape}")
        return np.ascontiguousarray(v)

    def nodes(self) -> np.ndarray:
        return np.asarray(self.mesh.centroids(), float)

    def collocation_points(self) -> np.ndarray:
        return np.ascontiguousarray(self.nodes())

    def collocation_normals(self) -> np.ndarray:
        normals, _ = self.mesh.normals_and_areas()
        return np.ascontiguousarray(np.asarray(normals, float))

    def collocation_shape(self) -> np.ndarray:
        return np.ones((1, 1))

    def value_array(self) -> np.ndarray:
        """The datum at the dof nodes, (n_dofs, 3) (core.py:355-402, P0)."""
        n = self.n_nodes
        if self.value is None:
            return np.zeros((n, 3))
        if callable(self.value):
            v = np.asarray(self.value(self.nodes()), dtype=float)
            if v.shape != (n, 3):
                raise ValueError(f"patch '{self.name}': callable value returned {v.shape}, "
                                 f"expected {(n, 3)}")
            return np.ascontiguousarray(v)
        v = np.asarray(self.value, dtype=float)
        return np.ascontiguousarray(np.broadcast_to(v, (n, 3)))

    def density(self) -> np.ndarray:
        return self.expand(self.value_array())

    def collocation_values(self, nodal) -> np.ndarray:
        return np.asarray(nodal, float).reshape(-1, 3)

    def integrate(self, values) -> np.ndarray:
        v = self.expand(values)
        _, areas = self.mesh.normals_and_areas()
        return (np.asarray(areas, float)[:, None] * v).sum(axis=0)


@dataclass(eq=False)
class Region:
    """A region bounded by oriented patches (core.py:500-513)."""
    name: str
    material: object
    patches: list
    probe_point: np.ndarray
    faults: list = field(default_factory=list)
    orientation_overrides: dict = field(default_factory=dict)

    @property
    def is_full_space(self) -> bool:
        return len(self.patches) == 0


def patch_solid_angle(patch: Patch, x) -> float:
    from ..geometry import solid_angle_batch
    return float(solid_angle_batch(patch.mesh, np.asarray(x, float))[0])


_CLOSED_THRESHOLD = 2.0 * np.pi
_OPEN_FLOOR = 1e-3
_CLOSURE_TOL = 1e-6


class RegionModel:
    """Regions + inferred orientations (core.py:702-1107, the P0 subset)."""

    def __init__(self, regions, gravity=None, far_field=None):
        if gravity is not None:
            raise NotImplementedError("gravity (buoyancy terms) is not in mbem_amd")
        if far_field is not None:
            raise NotImplementedError("far-field windows are not in mbem_amd")
        self.regions = list(regions)
        self.gravity = None
        self.far_field = None
        self._sigma: dict = {}
        self.validate()

    def _infer_sigma(self, region: Region, patch: Patch) -> int:
        """core.py:713-729."""
        if patch.name in region.orientation_overrides:
            return int(region.orientation_overrides[patch.name])
        omega = patch_solid_angle(patch, region.probe_point)
        if abs(omega) > _CLOSED_THRESHOLD or abs(omega) > _OPEN_FLOOR:
            return 1 if omega > 0 else -1
        cavity_probe = np.asarray(patch.mesh.vertices, float).mean(axis=0)
        omega_in = patch_solid_angle(patch, cavity_probe)
        if abs(omega_in) < _CLOSED_THRESHOLD:
            raise ValueError(
                f"cannot infer orientation of patch '{patch.name}' for region "
                f"'{region.name}': solid angle from probe is {omega:.3e} sr and the patch "
                f"does not enclose its own vertex centroid either; supply "
                f"orientation_overrides")
        return -1 if omega_in > 0 else 1

    def orientation(self, region: Region, patch: Patch) -> int:
        """sigma(region, patch): FAULT_ORIENTATION for a fault of the region
        (core.py:731-747)."""
        if patch.bc in FAULT_KINDS:
            if not any(f is patch for f in region.faults):
                raise ValueError(f"patch '{patch.name}' is a {patch.bc.name} but is not a fault "
                                 f"of region '{region.name}'")
            return FAULT_ORIENTATION
        return self._sigma[(region.name, patch.name)]

    def validate(self) -> None:
        """core.py:751-871: unique names, incidence, closure, interface
        antisymmetry, fault ownership and containment."""
        names = [r.name for r in self.regions]
        if len(set(names)) != len(names):
            raise ValueError("region names must be unique")
        incidence, by_id, pnames = {}, {}, {}
        for r in self.regions:
            for p in r.patches:
                incidence.setdefault(id(p), []).append(r)
                by_id[id(p)] = p
                existing = pnames.get(p.name)
                if existing is not None and existing is not p:
                    raise ValueError(f"two distinct Patch objects share the name '{p.name}' — "
                                     f"interface patches must be SHARED object instances")
                pnames[p.name] = p
        for pid, regs in incidence.items():
            p = by_id[pid]
            if p.bc is BCType.INTERFACE and len(regs) != 2:
                raise ValueError(f"interface patch '{p.name}' must belong to exactly 2 regions, "
                                 f"found {len(regs)}")
            if p.bc in (BCType.FREE_TRACTION, BCType.PRESCRIBED_DISPLACEMENT) and len(regs) != 1:
                raise ValueError(f"patch '{p.name}' ({p.bc.value}) must belong to exactly 1 "
                                 f"region, found {len(regs)}")
            if p.bc in FAULT_KINDS:
                raise ValueError(f"fault patch '{p.name}' belongs in Region.faults, not "
                                 f"Region.patches")
        for r in self.regions:
            if r.is_full_space:
                if not r.faults:
                    raise ValueError(f"region '{r.name}' has no boundary patch and no fault: "
                                     f"nothing to solve or evaluate")
                continue
            total = 0.0
            for p in r.patches:
                sigma = self._infer_sigma(r, p)
                self._sigma[(r.name, p.name)] = sigma
                total += sigma * patch_solid_angle(p, r.probe_point)
            if abs(total - 4.0 * np.pi) > _CLOSURE_TOL * 4.0 * np.pi:
                raise ValueError(
                    f"region '{r.name}' fails the closure identity: sum sigma*Omega = "
                    f"{total:.6f} sr, expected 4*pi = {4 * np.pi:.6f}. Check probe_point is "
                    f"interior and patch windings/orientation_overrides.")
        for pid, regs in incidence.items():
            p = by_id[pid]
            if p.bc is BCType.INTERFACE:
                s0 = self._sigma[(regs[0].name, p.name)]
                s1 = self._sigma[(regs[1].name, p.name)]
                if s0 != -s1:
                    raise ValueError(f"interface '{p.name}' has non-antisymmetric orientations: "
                                     f"{regs[0].name}:{s0}, {regs[1].name}:{s1}")
        fowner, fnames = {}, {}
        for r in self.regions:
            for f in r.faults:
                if f.bc not in FAULT_KINDS:
                    raise ValueError(f"fault '{f.name}' of region '{r.name}' has bc "
                                     f"{f.bc.value}; Region.faults holds FAULT patches only")
                fowner.setdefault(id(f), (f, []))[1].append(r.name)
                if f.name in pnames or fnames.get(f.name, f) is not f:
                    raise ValueError(f"fault '{f.name}' of region '{r.name}' shares its name "
                                     f"with another patch or fault; names must be unique")
                fnames[f.name] = f
        for f, owners in fowner.values():
            if len(owners) != 1:
                raise ValueError(f"fault '{f.name}' must belong to exactly 1 region, found "
                                 f"{len(owners)}: {owners}")
        from ..geometry import distance_to_mesh
        for r in self.regions:
            for f in r.faults:
                cs = f.mesh.centroids()
                for p in r.patches:
                    d, idx = distance_to_mesh(cs, p.mesh)
                    tri = np.asarray(p.mesh.triangles)[idx]
                    v = np.asarray(p.mesh.vertices, float)
                    h = np.linalg.norm(v[tri[:, 1]] - v[tri[:, 0]], axis=1)
                    k = int(np.argmin(d / h))
                    if d[k] < _ON_PATCH_TOL * h[k]:
                        raise ValueError(f"fault '{f.name}' triangle {k} lies ON boundary patch "
                                         f"'{p.name}' of region '{r.name}'")
                inside, total = self.point_in_region(r, cs, closure=True)
                if not inside.all():
                    k = int(np.argmax(~inside))
                    raise ValueError(f"fault '{f.name}' triangle {k} is not inside region "
                                     f"'{r.name}': closure sum at its centroid = "
                                     f"{total[k]:.6f} sr, expected 4*pi")

    def value_of(self, patch: Patch) -> np.ndarray:
        """The datum of ``patch`` at its dof nodes (core.py:893-903)."""
        return patch.value_array()

    def far_field_patches(self) -> list:
        return []

    def point_in_region(self, region, points, closure: bool = False):
        """Gauss-closure classifier (core.py:994-1026)."""
        from ..geometry import solid_angle_batch
        region = region if not isinstance(region, str) else \
            next(r for r in self.regions if r.name == region)
        pts = np.asarray(points, float).reshape(-1, 3)
        if region.is_full_space:
            total = np.full(pts.shape[0], 4.0 * np.pi)
            inside = np.ones(pts.shape[0], bool)
            return (inside, total) if closure else inside
        total = np.zeros(pts.shape[0])
        for p in region.patches:
            total += self._sigma[(region.name, p.name)] * solid_angle_batch(p.mesh, pts)
        inside = np.abs(total - 4.0 * np.pi) <= _CLOSURE_TOL * 4.0 * np.pi
        return (inside, total) if closure else inside

    def is_anchored(self) -> bool:
        """Some patch prescribes displacement (core.py:1028-1034)."""
        return any(p.bc is BCType.PRESCRIBED_DISPLACEMENT for r in self.regions
                   for p in r.patches)

    def interface_regions(self, patch: Patch) -> list:
        return [r for r in self.regions if any(q is patch for q in r.patches)]

    def slip_side(self, patch: Patch):
        return None

    def solved_faults(self) -> list:
        return []

    def sigma_table(self) -> dict:
        return dict(self._sigma)

    def with_materials(self, material_map: dict) -> "RegionModel":
        """This model with some regions' materials replaced; geometry and the
        orientation table shared (core.py:1081-1107)."""
        unknown = set(material_map) - {r.name for r in self.regions}
        if unknown:
            raise KeyError(f"unknown region(s) {sorted(unknown)} "
                           f"(regions: {[r.name for r in self.regions]})")
        new = RegionModel.__new__(RegionModel)
        new.regions = [replace(r, material=material_map[r.name]) if r.name in material_map
                       else r for r in self.regions]
        new.gravity = None
        new.far_field = None
        new._sigma = self._sigma
        return new
