"""Sample a solved model on a 3-D grid (mbem's src/mbem/volume.py): the
volumetric displacement, elastic stress, strain and eigenstress at grid
points inside each region, evaluated by the GPU readout
(``mbem_amd.evaluate``) with that point's own region's material and patches.

* ``make_grid`` / ``section_grid``: a regular grid (or one plane of it) over
  the model's bounding box, offset by half a cell so no plane of samples lies
  on an axis-aligned face (a point on a face has an ambiguous solid angle);
* ``classify``: 0 outside, else the 1-based region index from the model's
  Gauss-closure test, points ON a patch given to the first region owning it;
* ``clearance``: distance to the nearest boundary patch over its element
  size and over its eps, and to the nearest fault over its eps;
* ``stand_off``: with a ``recover.NearField``, a point inside
  ``standoff_h`` of an element of a recovered patch is read that far inside;
* ``sample``: the fields, chunk by chunk (NaN outside the body), and the
  derived scalars (|u|, von Mises, mean stress, max shear, dilatation);
* ``difference``, ``as_vti_arrays``: the difference of two samples with its
  scalars rebuilt, and the name -> (nz, ny, nx, ...) arrays ``vti.write`` takes.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import defaults
from .evaluate import evaluate_displacement, evaluate_stress
from .geometry import distance_to_mesh
from .kernels import basis as kb

SYM6 = (("xx", 0, 0), ("yy", 1, 1), ("zz", 2, 2), ("xy", 0, 1), ("yz", 1, 2), ("xz", 0, 2))
SCALARS = ("u_mag", "von_mises", "mean_stress", "max_shear", "dilatation")
PRIMARY = ("u", "sigma", "strain", "eigenstress")


@dataclasses.dataclass(frozen=True)
class Grid:
    """A regular grid of points, x fastest (ImageData order)."""
    origin: tuple
    spacing: tuple
    dims: tuple                      # (nx, ny, nz) point counts
    points: np.ndarray               # (N, 3), C order over (nz, ny, nx)

    @property
    def shape(self) -> tuple:
        nx, ny, nz = self.dims
        return (nz, ny, nx)

    def fold(self, a: np.ndarray) -> np.ndarray:
        """(N, ...) -> (nz, ny, nx, ...)."""
        return a.reshape(self.shape + a.shape[1:])


def model_bounds(model) -> tuple:
    """(lo, hi) of every boundary patch's vertices (volume.py:122-136)."""
    v = np.concatenate([np.asarray(p.mesh.vertices, float) for r in model.regions
                        for p in r.patches])
    return v.min(axis=0), v.max(axis=0)


def _axis(lo, hi, spacing):
    start = lo + 0.5 * spacing
    return start + spacing * np.arange(int(np.floor((hi - start) / spacing)) + 1)


def _grid(axes, spacing):
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return Grid(origin=tuple(float(a[0]) for a in axes), spacing=(float(spacing),) * 3,
                dims=tuple(len(a) for a in axes),
                points=np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]))


def make_grid(model, spacing: float, pad: float = 0.0) -> Grid:
    """The model's box widened by ``pad``, sampled every ``spacing`` from half a
    cell inside its low corner (volume.py:139-165)."""
    lo, hi = model_bounds(model)
    return _grid([_axis(lo[d] - pad, hi[d] + pad, spacing) for d in range(3)], spacing)


def section_grid(model, axis: str, at: float, spacing: float) -> Grid:
    """One plane ``axis = at`` of the grid ``make_grid`` lays out (volume.py:168-196);
    a plane outside the model is refused."""
    if axis not in ("x", "y", "z"):
        raise ValueError(f"a section is across x, y or z, not {axis!r}")
    k = "xyz".index(axis)
    lo, hi = model_bounds(model)
    if not lo[k] < at < hi[k]:
        raise ValueError(f"the section {axis} = {at} is outside the model, which spans "
                         f"{lo[k]:g} to {hi[k]:g} there")
    axes = [np.array([float(at)]) if d == k else _axis(lo[d], hi[d], spacing) for d in range(3)]
    return _grid(axes, spacing)


def on_patch(points: np.ndarray, patch) -> np.ndarray:
    """(N,) bool: the points within the model's on-patch tolerance of an
    element's size of ``patch`` (volume.py:199-227): the plane distance of
    the nearest elements first, the exact distance for the candidates."""
    from scipy.spatial import cKDTree

    from .model.core import _ON_PATCH_TOL
    mesh = patch.mesh
    tv = np.asarray(mesh.vertices, float)[np.asarray(mesh.triangles)]
    nrm = np.asarray(mesh.normals_and_areas()[0], float)
    tol = _ON_PATCH_TOL * kb.element_sizes(mesh)
    pts = np.asarray(points, float)
    on = np.zeros(pts.shape[0], bool)
    k = min(8, tv.shape[0])
    tree = cKDTree(tv.mean(axis=1))
    for s in range(0, pts.shape[0], defaults.VOLUME_CHUNK):
        p = pts[s:s + defaults.VOLUME_CHUNK]
        _, idx = tree.query(p, k=k)
        idx = np.asarray(idx).reshape(p.shape[0], k)
        plane = np.abs(np.einsum("nki,nki->nk", p[:, None, :] - tv[idx, 0], nrm[idx]))
        cand = np.flatnonzero((plane < tol[idx]).any(axis=1))
        if cand.size:
            d, j = distance_to_mesh(p[cand], mesh)
            on[s + cand[d < tol[j]]] = True
    return on


def classify(model, grid: Grid) -> np.ndarray:
    """(N,) int8: 0 outside every region, else the 1-based region index
    (volume.py:230-252)."""
    pts = grid.points
    code = np.zeros(pts.shape[0], np.int8)
    for i, r in enumerate(model.regions, start=1):
        code[model.point_in_region(r, pts) & (code == 0)] = i
    settled = np.zeros(code.shape[0], bool)
    for i, r in enumerate(model.regions, start=1):
        for p in r.patches:
            on = on_patch(pts, p) & ~settled
            code[on] = i
            settled |= on
    return code


def _patch_eps(eps, patch, idx):
    return np.broadcast_to(np.atleast_1d(kb.resolve_patch_eps(eps, patch)),
                           (patch.mesh.n_triangles,))[idx]


def clearance(model, grid: Grid, eps) -> tuple:
    """(clearance_h, clearance_eps, fault_eps), each (N,): distance to the
    nearest boundary patch over its element's size and over its eps, and to
    the nearest fault over its eps (volume.py:255-301)."""
    n = grid.points.shape[0]
    c_h, c_eps, c_f = (np.full(n, np.inf) for _ in range(3))
    for r in model.regions:
        for p in r.patches:
            d, idx = distance_to_mesh(grid.points, p.mesh)
            c_h = np.minimum(c_h, d / kb.element_sizes(p.mesh)[idx])
            c_eps = np.minimum(c_eps, d / _patch_eps(eps, p, idx))
        for f in r.faults:
            d, idx = distance_to_mesh(grid.points, f.mesh)
            c_f = np.minimum(c_f, d / _patch_eps(eps, f, idx))
    return c_h, c_eps, c_f


def strain_of(sig: np.ndarray, mu: float, lam: float) -> np.ndarray:
    """Hooke's law inverted from (mu, lam): tr e = tr sigma / (2 mu + 3 lam)."""
    tr_e = np.trace(sig, axis1=1, axis2=2) / (2.0 * mu + 3.0 * lam)
    return (sig - lam * tr_e[:, None, None] * np.eye(3)) / (2.0 * mu)


def derived(u: np.ndarray, sig: np.ndarray, e: np.ndarray) -> dict:
    """The scalars a volume render is coloured by (volume.py:317-328)."""
    tr = np.trace(sig, axis1=1, axis2=2)
    dev = sig - (tr / 3.0)[:, None, None] * np.eye(3)
    w = np.linalg.eigvalsh(sig)
    return {"u_mag": np.linalg.norm(u, axis=1),
            "von_mises": np.sqrt(1.5 * np.einsum("nij,nij->n", dev, dev)),
            "mean_stress": tr / 3.0, "max_shear": 0.5 * (w[:, 2] - w[:, 0]),
            "dilatation": np.trace(e, axis1=1, axis2=2)}


def stand_off(model, region, points: np.ndarray, eps, near) -> np.ndarray:
    """``points`` of ``region`` closer to a recovered patch than
    ``near.standoff_h`` of the nearest element's size, moved to that distance
    inward along the element's normal (volume.py:331-367)."""
    from scipy.spatial import cKDTree
    pts = np.array(points, float)
    best = np.ones(pts.shape[0])
    shift = np.zeros_like(pts)
    for p in region.patches:
        if not near.recovers(p):
            continue
        h = kb.element_sizes(p.mesh)
        gap = near.standoff_h * h
        d0, _ = cKDTree(p.mesh.centroids()).query(pts, k=1)
        cand = np.flatnonzero(d0 <= float((h + gap).max()))
        if cand.size == 0:
            continue
        d, idx = distance_to_mesh(pts[cand], p.mesh)
        ratio = d / gap[idx]
        take = ratio < best[cand]
        if not take.any():
            continue
        inward = -float(model.orientation(region, p)) * np.asarray(p.mesh.normals_and_areas()[0], float)
        who = cand[take]
        best[who] = ratio[take]
        shift[who] = (gap[idx] - d)[take, None] * inward[idx[take]]
    return pts + shift


def sample(model, solution: dict, eps, grid: Grid, code: np.ndarray, chunk: int = None,
           eigenstress: bool = True, progress=None, near=None) -> dict:
    """One solved state on ``grid`` (volume.py:370-444): (N, ...) arrays of
    u, sigma (elastic), strain, eigenstress (with ``eigenstress``) and the
    derived scalars, NaN where ``code`` is 0."""
    chunk = chunk or defaults.VOLUME_CHUNK
    n = grid.points.shape[0]
    if near is not None and not any(near.recovers(p) for r in model.regions for p in r.patches):
        raise ValueError("this NearField recovers no patch of the model: it was built for "
                         "another model, or the model has no boundary patch of order 0")
    out = {"u": np.full((n, 3), np.nan), "sigma": np.full((n, 3, 3), np.nan),
           "strain": np.full((n, 3, 3), np.nan)}
    if eigenstress:
        out["eigenstress"] = np.full((n, 3, 3), np.nan)
    for i, r in enumerate(model.regions, start=1):
        where = np.flatnonzero(code == i)
        mu, lam = r.material.mu, r.material.lam
        for s in range(0, where.size, chunk):
            sel = where[s:s + chunk]
            pts = grid.points[sel]
            if near is not None:
                pts = stand_off(model, r, pts, eps, near)
            out["u"][sel] = evaluate_displacement(model, r, solution, pts, eps, warn_near=False,
                                                  near=near)
            got = evaluate_stress(model, r, solution, pts, eps, subtract_anelastic=True,
                                  warn_near=False, parts=eigenstress, near=near)
            sig = got[0] if eigenstress else got
            out["sigma"][sel] = sig
            out["strain"][sel] = strain_of(sig, mu, lam)
            if eigenstress:
                out["eigenstress"][sel] = got[1]
            if progress is not None:
                progress(r.name, min(s + chunk, where.size), where.size)
    out.update(_scalars(out, code > 0))
    return out


def _scalars(fields, inside):
    sc = {k: np.full(inside.size, np.nan) for k in SCALARS}
    sel = np.flatnonzero(inside)
    if sel.size:
        for k, v in derived(fields["u"][sel], fields["sigma"][sel], fields["strain"][sel]).items():
            sc[k][sel] = v
    return sc


def difference(fa: dict, fb: dict, inside: np.ndarray) -> dict:
    """fa - fb on the vector and tensor fields, the scalars rebuilt from the
    difference (a von Mises of a difference is not a difference of von
    Mises values) (volume.py:447-471)."""
    out = {}
    for k in PRIMARY:
        if k in fa and k in fb:
            d = fa[k] - fb[k]
            d[~inside] = np.nan
            out[k] = d
    out.update(_scalars(out, inside))
    return out


def as_vti_arrays(grid: Grid, fields: dict, code: np.ndarray, c_h: np.ndarray,
                  c_eps: np.ndarray, c_fault: np.ndarray = None) -> dict:
    """name -> (nz, ny, nx, ...) for ``vti.write``: tensors as their six
    components, vectors and scalars as they are, plus the region code and
    the clearances (volume.py:474-497)."""
    arrays = {}
    for name, a in fields.items():
        if a.ndim == 3:
            for comp, i, j in SYM6:
                arrays[f"{name}_{comp}"] = grid.fold(a[:, i, j])
        else:
            arrays[name] = grid.fold(a)
    arrays["region"] = grid.fold(code.astype(float))
    arrays["clearance_h"] = grid.fold(c_h)
    arrays["clearance_eps"] = grid.fold(c_eps)
    if c_fault is not None:
        arrays["fault_eps"] = grid.fold(c_fault)
    return arrays
