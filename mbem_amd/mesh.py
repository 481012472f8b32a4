"""Triangle meshes, materials and the structured meshers of the box-and-fault
workloads.

``TriMesh`` / ``ElasticMaterial`` keep the reference types' fields and methods
(mbem's src/mollified_bem.py:31-44, :183-219).  The meshers follow
src/local_box_mesh.py: a tensor grid with the 4-fold mirror-symmetric
diagonal pattern (:38-91), flat patches (:94-108), vertical walls flipped to
an outward normal (:111-152), concatenation without vertex merging
(:222-238), and a vertical rectangular fault (:241-290) that
``tilted_fault`` turns about its trace as cases/layered_box.py:59-71 does.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class ElasticMaterial:
    """Isotropic elastic material (mollified_bem.py:31-44)."""
    mu: float
    lam: float

    @property
    def nu(self):
        return self.lam / (2.0 * (self.lam + self.mu))

    @property
    def E(self):
        return self.mu * (3.0 * self.lam + 2.0 * self.mu) / (self.lam + self.mu)


@dataclass
class TriMesh:
    """Triangle surface mesh (mollified_bem.py:183-219)."""
    vertices: np.ndarray
    triangles: np.ndarray

    @property
    def n_triangles(self):
        return self.triangles.shape[0]

    @property
    def n_vertices(self):
        return self.vertices.shape[0]

    def centroids(self):
        return self.vertices[self.triangles].mean(axis=1)

    def normals_and_areas(self):
        v = self.vertices[self.triangles]
        cross = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        area2 = np.linalg.norm(cross, axis=1)
        return cross / area2[:, None], 0.5 * area2

    def ensure_outward_normals(self, center=np.array([0.0, 0.0, 0.0])):
        normals, _ = self.normals_and_areas()
        flip = np.sum(normals * (self.centroids() - center), axis=1) < 0
        self.triangles[flip] = self.triangles[flip][:, [0, 2, 1]]


def rect_grid_triangles(nx: int, ny: int, flip: bool, symmetric: bool = True) -> np.ndarray:
    """Triangles of an (nx+1) x (ny+1) grid, vertex (i, j) = j (nx+1) + i
    (local_box_mesh.py:38-91): NW-SE diagonals in the (++) and (--)
    quadrants, NE-SW in the others, so the mesh is mirror symmetric."""
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    row = nx + 1
    v00 = j * row + i
    v10 = v00 + 1
    v01 = v00 + row
    v11 = v01 + 1
    if symmetric:
        nwse = ((2 * i + 1) < nx) == ((2 * j + 1) < ny)
    else:
        nwse = np.ones_like(i, dtype=bool)
    a = np.where(nwse[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    b = np.where(nwse[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    if flip:
        a = a[:, [0, 2, 1]]
        b = b[:, [0, 2, 1]]
    out = np.empty((2 * i.size, 3), dtype=np.int64)
    out[0::2] = a
    out[1::2] = b
    return out


def rectangular_patch(x_range, y_range, z_level, nx, ny, normal_up=True) -> TriMesh:
    """Flat rectangle at fixed z (local_box_mesh.py:94-108)."""
    xs = np.linspace(x_range[0], x_range[1], nx + 1)
    ys = np.linspace(y_range[0], y_range[1], ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    verts = np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, float(z_level))])
    return TriMesh(verts, rect_grid_triangles(nx, ny, flip=not normal_up))


def vertical_panel(axis, fixed, t_range, z_range, nt, nz, outward_sign) -> TriMesh:
    """One vertical wall of a box, normal along +-axis (local_box_mesh.py:111-152)."""
    if axis not in ("x", "y"):
        raise ValueError("axis must be 'x' or 'y'")
    ts = np.linspace(t_range[0], t_range[1], nt + 1)
    zs = np.linspace(z_range[0], z_range[1], nz + 1)
    T, Z = np.meshgrid(ts, zs, indexing="xy")
    if axis == "x":
        xs, ys = np.full(T.size, float(fixed)), T.ravel()
    else:
        xs, ys = T.ravel(), np.full(T.size, float(fixed))
    mesh = TriMesh(np.column_stack([xs, ys, Z.ravel()]), rect_grid_triangles(nt, nz, flip=False))
    normals, _ = mesh.normals_and_areas()
    want = np.zeros(3)
    want[0 if axis == "x" else 1] = float(outward_sign)
    if np.dot(normals[0], want) < 0:
        mesh.triangles = mesh.triangles[:, [0, 2, 1]]
    return mesh


def concatenate(meshes) -> TriMesh:
    """Meshes joined without vertex merging (local_box_mesh.py:222-238)."""
    verts, tris, off = [], [], 0
    for m in meshes:
        verts.append(m.vertices)
        tris.append(m.triangles + off)
        off += m.n_vertices
    return TriMesh(np.vstack(verts), np.vstack(tris))


def vertical_fault(strike_length, depth_range, n_along, n_down, center_xy=(0.0, 0.0),
                   strike_azimuth_deg=0.0):
    """Vertical rectangular fault striking at the azimuth (0 = +y); returns
    (mesh, n_hat, s_hat) with the stored normal along n_hat
    (local_box_mesh.py:241-290)."""
    z_bot, z_top = depth_range
    if not (z_bot < z_top <= 0):
        raise ValueError("expected z_bot < z_top <= 0 in depth_range")
    th = np.radians(strike_azimuth_deg)
    s_hat = np.array([np.sin(th), np.cos(th), 0.0])
    n_hat = np.array([np.cos(th), -np.sin(th), 0.0])
    ss = np.linspace(-strike_length / 2, strike_length / 2, n_along + 1)
    zs = np.linspace(z_bot, z_top, n_down + 1)
    Sg, Zg = np.meshgrid(ss, zs, indexing="xy")
    verts = np.column_stack([center_xy[0] + Sg.ravel() * s_hat[0],
                             center_xy[1] + Sg.ravel() * s_hat[1], Zg.ravel()])
    mesh = TriMesh(verts, rect_grid_triangles(n_along, n_down, flip=False))
    normals, _ = mesh.normals_and_areas()
    if np.dot(normals[0], n_hat) < 0:
        mesh.triangles = mesh.triangles[:, [0, 2, 1]]
    return mesh, n_hat, s_hat


def tilted_fault(fault: TriMesh, dip_deg: float):
    """The vertical panel turned about its trace (z = 0 line) to ``dip``
    degrees, dipping toward +x, and its normal (sin, 0, cos): up and into the
    hanging wall (cases/layered_box.py:59-71)."""
    d = np.deg2rad(dip_deg)
    R = np.array([[np.sin(d), 0.0, -np.cos(d)],
                  [0.0, 1.0, 0.0],
                  [np.cos(d), 0.0, np.sin(d)]])
    return TriMesh(fault.vertices @ R.T, fault.triangles.copy()), R[:, 0].copy()
