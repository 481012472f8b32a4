"""Manufactured boundary-value problems on a closed box, for pinning the
readout (the kind the reference's anchor B uses,
tests/gates/mbem/verify_solved_bvp.py:44-76, 907-935, 1188-1237, 1266-1290).

The box is the workload's 40 x 40 x 20 km box on 8 cells per 40 km
(h = 5 km, 512 triangles, 1,536 unknowns), small enough for the oracle to
assemble and LU-solve in about a second.  Units: km and GPa.

* ``rigid``: u = c on every face (all prescribed displacement).  Exact
  answer: t = 0 on the faces, u = c and sigma = 0 inside.
* ``strain``: u = E x on every face with E symmetric and traced, so lam
  enters the answer.  Exact answer: sigma = lam tr(E) I + 2 mu E inside.
* ``half``: the same strain field with the traction prescribed on the top
  and the walls (its exact, piecewise-constant value) and the displacement
  on the base.  A readout at a point ON the top patch returns half of the
  interior field, displacement and stress alike.
"""
from __future__ import annotations

import numpy as np

MU = 30.0
N_CELLS = 8
H = 40.0 / N_CELLS
EPS = 0.125 * H                      # eps/h = 1/8 on every face
RIGID_C = np.array([0.004, -0.0015, 0.0025])
# a symmetric strain with a trace (lam is live), not aligned with the box
STRAIN = 1.0e-4 * np.array([[1.5, -0.5, 0.8],
                            [-0.5, 2.0, -1.0],
                            [0.8, -1.0, -0.5]])
HALF_BARY = np.array([0.6, 0.25, 0.15])    # not the centroid (a collocation point)


def material(nu, mu=MU):
    from mbem_amd.mesh import ElasticMaterial
    return ElasticMaterial(mu=mu, lam=2.0 * mu * nu / (1.0 - 2.0 * nu))


def exact_sigma(mat, E=STRAIN):
    return mat.lam * np.trace(E) * np.eye(3) + 2.0 * mat.mu * E


def _meshes():
    from mbem_amd.cases import box_meshes
    top, sides, base = box_meshes(N_CELLS)
    return {"top": top, "sides": sides, "base": base}


def _model(mat, bcs, values):
    from mbem_amd.model import BCType, Patch, Region, RegionModel, generate_system
    meshes = _meshes()
    kind = {"u": BCType.PRESCRIBED_DISPLACEMENT, "t": BCType.FREE_TRACTION}
    patches = [Patch(k, meshes[k], kind[bcs[k]], value=values(k, meshes[k], bcs[k]))
               for k in ("top", "sides", "base")]
    region = Region("block", mat, patches, probe_point=np.array([1.0, 2.0, -10.0]))
    model = RegionModel([region])
    return model, region, generate_system(model)


def rigid(nu):
    """(model, region, system) of the rigid translation."""
    return _model(material(nu), dict(top="u", sides="u", base="u"),
                  lambda k, m, bc: np.broadcast_to(RIGID_C, (m.n_triangles, 3)).copy())


def _strain_value(mat, E):
    def value(k, mesh, bc):
        if bc == "u":
            return mesh.centroids() @ E.T
        n, _ = mesh.normals_and_areas()
        return n @ exact_sigma(mat, E).T
    return value


def strain(nu, E=STRAIN, mat=None):
    """(model, region, system) of the uniform strain, displacement on every
    face; ``mat`` overrides the model's material (the data stay those of
    ``nu``), which is how a material error is planted."""
    true = material(nu)
    return _model(mat or true, dict(top="u", sides="u", base="u"), _strain_value(true, E))


def half(nu, E=STRAIN):
    """(model, region, system) of the uniform strain with traction on the top
    and the walls, displacement on the base."""
    mat = material(nu)
    return _model(mat, dict(top="t", sides="t", base="u"), _strain_value(mat, E))


def interior_points(n=40, seed=5):
    """Points well inside the box: |x|, |y| <= 11 km, z in [-15, -5] km."""
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-11.0, 11.0, n), rng.uniform(-11.0, 11.0, n),
                            rng.uniform(-15.0, -5.0, n)])


def on_top_points(top_mesh):
    """One point per top triangle at HALF_BARY, kept to the middle of the
    patch (|x|, |y| < 12 km) so the box rim plays no part."""
    v = np.asarray(top_mesh.vertices, float)[np.asarray(top_mesh.triangles)]
    pts = np.einsum("tki,k->ti", v, HALF_BARY)
    keep = (np.abs(pts[:, 0]) < 12.0) & (np.abs(pts[:, 1]) < 12.0)
    pts = pts[keep]
    pts[:, 2] = 0.0
    return pts


def projections(sig, u, q, mat, E=STRAIN):
    """(stress, displacement) projections of the readout onto the exact
    fields: 1 in the interior, 1/2 on a patch (a projection, since u = E x
    is small near the origin)."""
    sx = np.broadcast_to(exact_sigma(mat, E), sig.shape)
    ue = q @ E.T
    return (float(np.sum(sig * sx) / np.sum(sx * sx)), float(np.sum(u * ue) / np.sum(ue * ue)))


def hydrostatic_error(sig, mat, E=STRAIN):
    """|median tr(sigma)/3 over the points / (tr(C:E)/3) - 1|."""
    ref = float(np.trace(exact_sigma(mat, E))) / 3.0
    return abs(float(np.median(np.trace(sig, axis1=1, axis2=2) / 3.0)) / ref - 1.0)


def tensor_error(sig, mat, E=STRAIN):
    """max over points of |sigma - C:E|_F / |C:E|_F."""
    sx = exact_sigma(mat, E)
    return float(np.max(np.linalg.norm(sig - sx, axis=(1, 2))) / np.linalg.norm(sx))


def oracle_solution(system, eps=EPS, jump="calibrated"):
    """The oracle's assembly and LU solution as the slot dict."""
    from oracle import dense as od
    A, b, _ = od.assemble(system, eps, jump=jump)
    return od.split(system.layout, od.lu_solve(A, b))


def exact_density_model(nu, n_cells=32):
    """(model, region, solution, mat) carrying the EXACT densities of the
    field u = c + E x (c = RIGID_C, E = STRAIN) on a box of ``n_cells``:
    u prescribed on every face at the element centroids, and the traction
    sigma n (constant per face) given as the solved t-slots.  No solve: the
    representation formula alone must give the field inside, half of it on a
    face and zero outside (Somigliana's identity), whatever the readout's
    internal signs -- which is what it pins."""
    from mbem_amd.cases import box_meshes
    from mbem_amd.model import BCType, Patch, Region, RegionModel
    mat = material(nu)
    top, sides, base = box_meshes(n_cells)
    meshes = {"top": top, "sides": sides, "base": base}
    patches = [Patch(k, m, BCType.PRESCRIBED_DISPLACEMENT, value=RIGID_C + m.centroids() @ STRAIN.T)
               for k, m in meshes.items()]
    region = Region("block", mat, patches, probe_point=np.array([1.0, 2.0, -10.0]))
    solution = {}
    for k, m in meshes.items():
        n, _ = m.normals_and_areas()
        solution[f"t:{k}"] = n @ exact_sigma(mat).T
    return RegionModel([region]), region, solution, mat


def outside_points(n=20, seed=7):
    """Points above the top face (z in [2, 6] km) and beyond a wall."""
    rng = np.random.default_rng(seed)
    up = np.column_stack([rng.uniform(-10.0, 10.0, n), rng.uniform(-10.0, 10.0, n),
                          rng.uniform(2.0, 6.0, n)])
    side = np.column_stack([rng.uniform(24.0, 28.0, n), rng.uniform(-10.0, 10.0, n),
                            rng.uniform(-15.0, -5.0, n)])
    return np.concatenate([up, side])
