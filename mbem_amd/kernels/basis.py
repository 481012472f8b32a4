"""Material coefficients, eps resolution and the assembly entry points
(mbem's src/mbem/kernels/basis.py, order 0).

The influence matrices decompose as ``M(material) = sum_k c_k(mu, lam) B_k``
with geometry-only basis blocks (3 for U, 6 for T).  Coefficients come from
(mu, lam) directly, never through 1/(1-2nu) (basis.py:14-19).  eps is an
(N_src,) per-element array everywhere; a scalar is promoted.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .. import defaults


def _mu_lam(material=None, mu=None, lam=None):
    if material is not None:
        return material.mu, material.lam
    if mu is None or lam is None:
        raise ValueError("provide either material or (mu, lam)")
    return mu, lam


def u_coeffs(mu, lam) -> np.ndarray:
    """(3,) coefficients of the U basis [I1 d, eps^2 I3 d, T2[3]] (basis.py:46-55)."""
    nu = lam / (2.0 * (lam + mu))
    one_minus_nu = (lam + 2.0 * mu) / (2.0 * (lam + mu))
    C1 = 1.0 / (16.0 * np.pi * mu * one_minus_nu)
    c34 = 3.0 - 4.0 * nu
    return np.array([c34 * C1, 1.0 / (8.0 * np.pi * mu), C1])


def t_coeffs(mu, lam) -> np.ndarray:
    """(6,) coefficients of the T basis [N1, N2, N3, R1, R2, R3] (basis.py:58-72):
    the N (trace) blocks carry lam*C1, the R blocks mu*C1."""
    nu = lam / (2.0 * (lam + mu))
    one_minus_nu = (lam + 2.0 * mu) / (2.0 * (lam + mu))
    C1 = 1.0 / (16.0 * np.pi * mu * one_minus_nu)
    c34 = 3.0 - 4.0 * nu
    cL = lam * C1
    cM = mu * C1
    blob = 6.0 * one_minus_nu
    return np.array([c34 * cL, -cL, blob * cL, c34 * cM, -cM, blob * cM])


def as_eps_array(eps, n_source: int) -> np.ndarray:
    """Promote a scalar eps to (N_src,), validate length and positivity
    (basis.py:187-204): every entry finite and > 0, else ValueError."""
    arr = np.asarray(eps, dtype=float)
    if arr.ndim == 0:
        arr = np.full(n_source, float(arr))
    elif arr.shape != (n_source,):
        raise ValueError(f"eps array shape {arr.shape} != ({n_source},)")
    bad = ~(np.isfinite(arr) & (arr > 0.0))
    if bad.any():
        raise ValueError(f"eps must be finite and > 0: {int(bad.sum())} bad "
                         f"of {arr.size}, min = {arr.min()}")
    return np.ascontiguousarray(arr)


def element_sizes(mesh) -> np.ndarray:
    """Per-triangle size h_j = mean edge length (basis.py:207-212)."""
    tv = np.asarray(mesh.vertices, float)[np.asarray(mesh.triangles)]
    e = np.stack([tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 1], tv[:, 0] - tv[:, 2]])
    return np.linalg.norm(e, axis=2).mean(axis=0)


def resolve_eps(eps, mesh) -> np.ndarray:
    """eps spec of one source mesh -> (N_src,) (basis.py:215-231): a scalar,
    an (N_src,) array, or "auto" = EPS_OVER_H * h_j."""
    if isinstance(eps, str):
        if eps != "auto":
            raise ValueError(f"unknown eps spec {eps!r} (only 'auto')")
        return np.ascontiguousarray(defaults.EPS_OVER_H * element_sizes(mesh))
    return as_eps_array(eps, mesh.n_triangles)


def resolve_patch_eps(spec, patch) -> np.ndarray:
    """eps of one patch or fault (basis.py:234-254): a dict is keyed by patch
    name; "auto" on a fault is one value FAULT_EPS_OVER_H * min(h)."""
    from ..model.core import FAULT_KINDS

    if isinstance(spec, dict):
        if patch.name not in spec:
            raise ValueError(f"eps dict has no entry for patch "
                             f"{patch.name!r} (keys: {sorted(spec)})")
        spec = spec[patch.name]
    if isinstance(spec, str) and spec == "auto" and patch.bc in FAULT_KINDS:
        h = element_sizes(patch.mesh)
        return np.full(h.size, defaults.FAULT_EPS_OVER_H * float(h.min()))
    return resolve_eps(spec, patch.mesh)


def _source_arrays(mesh_source):
    tri_verts = np.asarray(mesh_source.vertices, float)[np.asarray(mesh_source.triangles)]
    normals, _ = mesh_source.normals_and_areas()
    return np.ascontiguousarray(tri_verts), np.ascontiguousarray(normals)


def _source_geometry(mesh_source):
    """(tri_verts, normals, centroids) of a triangle mesh in one host loop of
    the library (mbem_tri_geometry): bit for bit what _source_arrays and
    mesh.centroids() give (the same FP64 operations in the same order)."""
    from .. import _lib
    verts = np.ascontiguousarray(mesh_source.vertices, dtype=np.float64)
    tris = np.ascontiguousarray(mesh_source.triangles, dtype=np.int64)
    if verts.ndim != 2 or verts.shape[1] != 3 or tris.ndim != 2 or tris.shape[1] != 3:
        raise ValueError(f"mesh arrays must be (nv, 3) and (nt, 3), got {verts.shape} and "
                         f"{tris.shape}")
    nt = tris.shape[0]
    tv = np.empty((nt, 3, 3))
    nrm = np.empty((nt, 3))
    cen = np.empty((nt, 3))
    _lib.call("mbem_tri_geometry", _lib.host_ptr(verts), verts.shape[0], _lib.host_ptr(tris), nt,
              _lib.host_ptr(tv), _lib.host_ptr(nrm), _lib.host_ptr(cen))
    return tv, nrm, cen


def _field_points(mesh_or_points):
    if hasattr(mesh_or_points, "centroids"):
        return np.ascontiguousarray(mesh_or_points.centroids())
    return np.ascontiguousarray(np.asarray(mesh_or_points, dtype=float))


class MeshArrays:
    """Memo of mesh-derived arrays for one assembly or readout (basis.py:91-150):
    source vertices and normals, field points, collocation points and normals,
    each computed once per object.  Keyed by ``id()``, so it must not outlive
    the build it serves (a mesh can be mutated in place, an id reused)."""

    def __init__(self):
        self._memo: dict = {}

    def _get(self, key, make):
        v = self._memo.get(key)
        if v is None:
            v = make()
            self._memo[key] = v
        return v

    def source_arrays(self, mesh_source):
        return self._get(("src", id(mesh_source)), lambda: _source_arrays(mesh_source))

    def field_points(self, mesh_or_points):
        if not hasattr(mesh_or_points, "centroids"):
            return _field_points(mesh_or_points)
        return self._get(("pts", id(mesh_or_points)), lambda: _field_points(mesh_or_points))

    def collocation_points(self, patch):
        return self._get(("coll", id(patch)),
                         lambda: np.ascontiguousarray(patch.collocation_points()))

    def collocation_normals(self, patch):
        return self._get(("nrm", id(patch)),
                         lambda: np.ascontiguousarray(patch.collocation_normals()))


# Lagrange nodes of a triangle in the nodal kernels' order (basis.py:152-185):
# order 0 the centroid; 1 the vertices; 2 the vertices, then the midpoints of
# edges 12, 23, 31.  Node k of element s is density row K s + k.

def lagrange_nodes(order: int, shrink: float = 0.0) -> np.ndarray:
    """Barycentric coordinates (K, 3) of the nodes, each pulled toward the
    centroid by ``shrink`` t: (1 - t) lam + t / 3 (a collocation point)."""
    from .tri_nodal import n_nodes
    K = n_nodes(order)
    if order == 0:
        lam = np.full((1, 3), 1.0 / 3.0)
    else:
        lam = np.eye(3)
        if order == 2:
            lam = np.vstack([lam, 0.5 * (lam + np.roll(lam, -1, axis=0))])
    t = float(shrink)
    if t:
        lam = (1.0 - t) * lam + t / 3.0
    assert lam.shape == (K, 3)
    return lam


def lagrange_shape(order: int, lam) -> np.ndarray:
    """Shape functions (M, K) at barycentric points ``lam`` (M, 3): 1; the
    barycentrics; l_i (2 l_i - 1) at the vertices and 4 l_i l_j at the edge
    midpoints."""
    from .tri_nodal import n_nodes
    lam = np.atleast_2d(np.asarray(lam, float))
    if n_nodes(order) == 1:
        return np.ones((lam.shape[0], 1))
    if order == 1:
        return lam.copy()
    nxt = np.roll(lam, -1, axis=1)
    return np.hstack([lam * (2.0 * lam - 1.0), 4.0 * lam * nxt])


def identity_ids(order: int, n_source: int) -> np.ndarray:
    """The element-major layout as an ids table (N_s, K): ids[s, k] = K s + k."""
    from .tri_nodal import n_nodes
    K = n_nodes(order)
    return np.arange(K * int(n_source), dtype=np.int64).reshape(int(n_source), K)


def condense_columns(blk, ids, n_cols: int, out=None) -> np.ndarray:
    """Element-major columns (R, 3 K N_s) summed into shared node columns
    (R, 3 n_cols): column 3 (K s + k) + j is ADDED to column 3 ids[s, k] + j
    (basis.py:283-301); ``out`` accumulates in place when given."""
    blk = np.asarray(blk, float)
    ids = np.asarray(ids, dtype=np.int64)
    N_s, K = ids.shape
    if blk.ndim != 2 or blk.shape[1] != 3 * K * N_s:
        raise ValueError(f"block has {blk.shape[-1]} columns, ids imply {3 * K * N_s}")
    shape = (blk.shape[0], 3 * int(n_cols))
    if out is None:
        out = np.zeros(shape)
    elif out.shape != shape:
        raise ValueError(f"out shape {out.shape} != {shape}")
    if ids.size and (ids.min() < 0 or ids.max() >= n_cols):
        raise ValueError(f"ids must index 0 .. n_cols - 1 = {int(n_cols) - 1}")
    dst = (3 * ids.reshape(-1, 1) + np.arange(3)).ravel()
    np.add.at(out.T, dst, blk.T)
    return out


@dataclass
class UBasis:
    """Geometry-only U basis stack (3, 3N_f, 3N_s) (basis.py:304-314)."""
    stack: np.ndarray

    def combine(self, material=None, mu=None, lam=None) -> np.ndarray:
        c = u_coeffs(*_mu_lam(material, mu, lam))
        return np.tensordot(c, self.stack, axes=1)

    def nbytes(self) -> int:
        return self.stack.nbytes


@dataclass
class TBasis:
    """Geometry-only T basis stack (6, 3N_f, 3N_s) (basis.py:317-327)."""
    stack: np.ndarray

    def combine(self, material=None, mu=None, lam=None) -> np.ndarray:
        c = t_coeffs(*_mu_lam(material, mu, lam))
        return np.tensordot(c, self.stack, axes=1)

    def nbytes(self) -> int:
        return self.stack.nbytes


def _condensed(kernel_call, x_field, tri_verts, normals, eps_arr, order, ids, n_cols):
    """``kernel_call`` on chunks of ASSEMBLY_SOURCE_CHUNK source elements, each
    chunk's element-major columns summed into the shared columns before the
    next (basis.py:354-369)."""
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
    from .tri_nodal import n_nodes
    N_s = tri_verts.shape[0]
    if ids.shape != (N_s, n_nodes(order)):
        raise ValueError(f"ids shape {ids.shape} != ({N_s}, {n_nodes(order)})")
    out = np.zeros((3 * x_field.shape[0], 3 * int(n_cols)))
    step = int(defaults.ASSEMBLY_SOURCE_CHUNK)
    for c0 in range(0, N_s, step):
        c1 = min(N_s, c0 + step)
        blk = kernel_call(x_field, tri_verts[c0:c1], None if normals is None else normals[c0:c1],
                          eps_arr[c0:c1], order)
        condense_columns(blk, ids[c0:c1], n_cols, out=out)
    return out


def assemble_u_basis(mesh_field, mesh_source, eps, arrays=None, order: int = 0) -> UBasis:
    """basis.py:336-342 (GPU kernels; P0 pair code at order 0, nodal beyond)."""
    from . import tri_nodal as tn
    x_field = _field_points(mesh_field)
    tri_verts, _ = _source_arrays(mesh_source)
    return UBasis(tn.u_basis_matrices(x_field, tri_verts, as_eps_array(eps, tri_verts.shape[0]),
                                      order))


def assemble_t_basis(mesh_field, mesh_source, eps, arrays=None, order: int = 0) -> TBasis:
    """basis.py:345-351 (GPU kernels; P0 pair code at order 0, nodal beyond)."""
    from . import tri_nodal as tn
    x_field = _field_points(mesh_field)
    tri_verts, normals = _source_arrays(mesh_source)
    return TBasis(tn.t_basis_matrices(x_field, tri_verts, normals,
                                      as_eps_array(eps, tri_verts.shape[0]), order))


def assemble_u_matrix(mesh_field, mesh_source, material, eps, arrays=None, order: int = 0,
                      ids=None, n_cols=None) -> np.ndarray:
    """One-shot U matrix (3 N_f, 3 K N_s), or (3 N_f, 3 n_cols) condensed onto
    shared-node columns when ``ids`` (N_s, K) is given (basis.py:372-386)."""
    from . import tri_nodal as tn
    x_field = _field_points(mesh_field)
    tri_verts, _ = _source_arrays(mesh_source)
    eps_arr = as_eps_array(eps, tri_verts.shape[0])
    g1, g2, g3 = u_coeffs(material.mu, material.lam)
    if ids is None:
        return tn.u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3, order)
    return _condensed(lambda x, tv, _n, e, p: tn.u_matrix_direct(x, tv, e, g1, g2, g3, p),
                      x_field, tri_verts, None, eps_arr, order, ids, n_cols)


def assemble_t_matrix(mesh_field, mesh_source, material, eps, arrays=None, order: int = 0,
                      ids=None, n_cols=None) -> np.ndarray:
    """One-shot T matrix (3 N_f, 3 K N_s), or condensed onto shared-node
    columns when ``ids`` is given (basis.py:389-403)."""
    from . import tri_nodal as tn
    x_field = _field_points(mesh_field)
    tri_verts, normals = _source_arrays(mesh_source)
    eps_arr = as_eps_array(eps, tri_verts.shape[0])
    c = t_coeffs(material.mu, material.lam)
    if ids is None:
        return tn.t_matrix_direct(x_field, tri_verts, normals, eps_arr, *c, order)
    return _condensed(lambda x, tv, n, e, p: tn.t_matrix_direct(x, tv, n, e, *c, p),
                      x_field, tri_verts, normals, eps_arr, order, ids, n_cols)


def assemble_traction_matrix(x_field, n_field, mesh_source, material, eps, arrays=None,
                             order: int = 0, kind: str = "slip", ids=None, n_cols=None):
    """Traction rows (3 N_f, 3 n_cols) at the points ``x_field`` with normals
    ``n_field`` of a nodal density on ``mesh_source`` (basis.py:406-427):
    ``kind="slip"`` the ELASTIC traction of a displacement jump (eigenstress
    removed: the rows of a traction-BC fault), ``"force"`` the Kelvin single
    layer's, ``"eigen"`` / ``"total"`` the eigenstress alone / left in.
    Element-major columns unless ``ids`` (N_s, K) maps them onto ``n_cols``
    shared nodes; the scatter happens inside the kernel."""
    from . import tri_nodal as tn
    tri_verts, normals = _source_arrays(mesh_source)
    eps_arr = as_eps_array(eps, tri_verts.shape[0])
    if ids is None:
        ids = identity_ids(order, tri_verts.shape[0])
        n_cols = ids.size
    return tn.traction_matrix(np.ascontiguousarray(np.asarray(x_field, float)),
                              np.ascontiguousarray(np.asarray(n_field, float)),
                              tri_verts, normals, eps_arr, ids, int(n_cols),
                              material.mu, material.lam, order, kind)
