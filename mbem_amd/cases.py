"""The box-with-dipping-fault workloads of the dense path.

A 40 x 40 x 20 km box (x, y in [-20, 20], z in [-20, 0]) meshed on a
structured grid with ``n_cells`` cells per 40 km on every face (the
symmetric-diagonal grid of src/local_box_mesh.py:38-91): top and base
2 n^2 triangles each, each wall n x n/2 cells.  Top: FREE_TRACTION (zero);
sides and base: PRESCRIBED_DISPLACEMENT (zero), so the model is anchored.
A 10 x 10 km planar fault, meshed vertically (local_box_mesh.py:241-290)
and turned about its trace to the dip (cases/layered_box.py:59-71), centred
at x = y = 0 with its top edge at z = -3 km, carries uniform unit reverse
slip: the Burgers vector is the up-dip unit vector (the hanging wall, on the
+n side, moves up dip).  eps = 0.5 x the patch's mean edge length, one value
per patch.  Units: km, GPa (mu = 30 GPa = 3e10 Pa, nu = 0.25 -> lam = 30).

n_cells = 16 gives N = 2,048 box triangles (n = 6,144 unknowns), n_cells =
50 gives N = 20,000 (n = 60,000; the FP64 matrix is 28.8 GB), n_cells = 36
gives N = 10,368 (n = 31,104).  ``random_box_fault`` draws the family the
learned preconditioner is trained on: fault strike, dip, size and the box
resolution varied.
"""
from __future__ import annotations

import numpy as np

from .kernels.basis import element_sizes
from .mesh import (ElasticMaterial, concatenate, rectangular_patch, tilted_fault,
                   vertical_fault, vertical_panel)
from .model import BCType, Patch, Region, RegionModel


def box_meshes(n_cells: int, half=20.0, depth=20.0):
    nz = max(1, n_cells // 2)
    xr = (-half, half)
    top = rectangular_patch(xr, xr, 0.0, n_cells, n_cells, normal_up=True)
    base = rectangular_patch(xr, xr, -depth, n_cells, n_cells, normal_up=False)
    sides = concatenate([
        vertical_panel("x", half, xr, (-depth, 0.0), n_cells, nz, +1),
        vertical_panel("x", -half, xr, (-depth, 0.0), n_cells, nz, -1),
        vertical_panel("y", half, xr, (-depth, 0.0), n_cells, nz, +1),
        vertical_panel("y", -half, xr, (-depth, 0.0), n_cells, nz, -1)])
    return top, sides, base


def dipping_fault(size=10.0, cells=(10, 20), dip=45.0, top_z=-3.0, strike=0.0):
    """(mesh, normal, burgers): the fault plane dipping toward +x, then turned
    by ``strike`` degrees about the vertical through the origin."""
    vert, _, _ = vertical_fault(size, (-size, 0.0), cells[0], cells[1])
    mesh, normal = tilted_fault(vert, dip)
    v = mesh.vertices.copy()
    xs = v[:, 0]
    v[:, 0] -= 0.5 * (xs.min() + xs.max())
    v[:, 2] += top_z - v[:, 2].max()
    d = np.deg2rad(dip)
    burgers = np.array([-np.cos(d), 0.0, np.sin(d)])
    if strike:
        a = np.deg2rad(strike)
        R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
        v = v @ R.T
        normal = R @ np.asarray(normal, float)
        burgers = R @ burgers
    mesh.vertices = v
    return mesh, normal, burgers


def box_fault_model(n_cells: int = 16, fault_cells=(10, 20), dip: float = 45.0,
                    mu: float = 30.0, lam: float = 30.0, eps_over_h: float = 0.5,
                    slip: float = 1.0, fault: bool = True, strike: float = 0.0,
                    fault_size: float = 10.0):
    """(model, eps spec, info) of the box-with-fault workload."""
    top, sides, base = box_meshes(n_cells)
    p_top = Patch("top", top, BCType.FREE_TRACTION)
    p_sides = Patch("sides", sides, BCType.PRESCRIBED_DISPLACEMENT)
    p_base = Patch("base", base, BCType.PRESCRIBED_DISPLACEMENT)
    faults = []
    eps = {}
    for p in (p_top, p_sides, p_base):
        eps[p.name] = eps_over_h * float(element_sizes(p.mesh).mean())
    info = {"n_cells": n_cells, "box_triangles": top.n_triangles + sides.n_triangles
            + base.n_triangles}
    if fault:
        fmesh, fnormal, burgers = dipping_fault(size=fault_size, cells=fault_cells, dip=dip,
                                                strike=strike)
        p_fault = Patch("fault", fmesh, BCType.FAULT,
                        value=np.broadcast_to(slip * burgers, (fmesh.n_triangles, 3)))
        faults.append(p_fault)
        eps["fault"] = eps_over_h * float(element_sizes(fmesh).mean())
        info.update(fault_triangles=fmesh.n_triangles, fault_normal=fnormal.tolist(),
                    burgers=burgers.tolist())
    region = Region("crust", ElasticMaterial(mu=mu, lam=lam), [p_top, p_sides, p_base],
                    probe_point=np.array([9.0, 9.0, -10.0]), faults=faults)
    return RegionModel([region]), eps, info


def random_box_fault_params(rng, n_cells_range=(26, 50)) -> dict:
    """The draw of ``random_box_fault`` alone (n_cells, strike, dip, size),
    so a caller can step a generator past geometries without building them."""
    lo, hi = n_cells_range
    n_cells = int(rng.integers(lo // 2, hi // 2 + 1)) * 2
    strike = float(rng.uniform(0.0, 180.0))
    dip = float(rng.uniform(20.0, 90.0))
    size = float(rng.uniform(5.0, 15.0))
    return {"n_cells": n_cells, "strike": strike, "dip": dip, "size": size}


def random_box_fault(rng, n_cells_range=(26, 50), params: dict = None):
    """(model, eps, info) of one random member of the training family: box
    resolution n_cells even in ``n_cells_range`` (N = 8 n^2 triangles: 5.4k
    to 20k for the default), fault strike U(0, 180) deg, dip U(20, 90) deg,
    size U(5, 15) km with cells matched to the box spacing.  ``params`` (a
    ``random_box_fault_params`` draw) builds that member instead of drawing."""
    prm = params if params is not None else random_box_fault_params(rng, n_cells_range)
    n_cells, strike, dip, size = prm["n_cells"], prm["strike"], prm["dip"], prm["size"]
    h = 40.0 / n_cells
    nf = max(2, int(round(size / h)))
    model, eps, info = box_fault_model(n_cells, fault_cells=(nf, 2 * nf), dip=dip, strike=strike,
                                       fault_size=size)
    info.update(strike=strike, dip=dip, size=size)
    return model, eps, info


def surface_grid(n=50, half=20.0, z=0.0):
    """An n x n grid of points over [-half, half]^2 at height z."""
    g = np.linspace(-half, half, n)
    X, Y = np.meshgrid(g, g, indexing="xy")
    return np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, z)])


def readout_case(n_points: int, seed_pts: int = 11):
    """(points, tri_verts, normals, eps, densities, info) of the stress-readout
    throughput case.  Sources: the N = 20,000 box + 676-triangle fault mesh,
    per-element eps = 0.5 h, random unit densities (seed 3).  Points: half
    uniform in the box interior away from the faces (0.1 h), half within
    3 eps of the fault plane (uniform over the fault, normal offset
    U(-3 eps, 3 eps))."""
    from .kernels import basis as kb
    model, _, info = box_fault_model(50, fault_cells=(13, 26))
    region = model.regions[0]
    patches = list(region.patches) + list(region.faults)
    tv, nrm, eps = [], [], []
    for p in patches:
        a, b = kb._source_arrays(p.mesh)
        tv.append(a)
        nrm.append(b)
        eps.append(0.5 * kb.element_sizes(p.mesh))
    tv, nrm, eps = np.concatenate(tv), np.concatenate(nrm), np.concatenate(eps)
    dens = np.random.default_rng(3).normal(size=(tv.shape[0], 3))
    dens /= np.linalg.norm(dens, axis=1, keepdims=True)
    rng = np.random.default_rng(seed_pts)
    h = 40.0 / 50
    nb = n_points // 2
    lo = np.array([-20.0, -20.0, -20.0]) + 0.1 * h
    hi = np.array([20.0, 20.0, 0.0]) - 0.1 * h
    box_pts = lo + (hi - lo) * rng.random((nb, 3))
    fault = region.faults[0]
    fv = fault.mesh.vertices[fault.mesh.triangles]
    area = 0.5 * np.linalg.norm(np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0]), axis=1)
    nf = n_points - nb
    tri = rng.choice(fv.shape[0], size=nf, p=area / area.sum())
    u, v = rng.random(nf), rng.random(nf)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    base = fv[tri, 0] + u[:, None] * (fv[tri, 1] - fv[tri, 0]) + v[:, None] * (fv[tri, 2] - fv[tri, 0])
    fn_ = np.cross(fv[0, 1] - fv[0, 0], fv[0, 2] - fv[0, 0])
    fn_ /= np.linalg.norm(fn_)
    eps_f = 0.5 * float(kb.element_sizes(fault.mesh).mean())
    fault_pts = base + rng.uniform(-3 * eps_f, 3 * eps_f, size=(nf, 1)) * fn_[None, :]
    pts = np.concatenate([box_pts, fault_pts])
    return pts, tv, nrm, eps, dens, info
