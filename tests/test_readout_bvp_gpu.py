"""The GPU readout (mbem_amd.evaluate) on the manufactured problems that pin
the oracle's readout (tests/test_readout_pinning.py), through the product
path: device assembly, FGMRES, and the device contractions.

* Somigliana's identity with exact densities: the same bars as the oracle
  (interior stress 1e-3, displacement 5e-4; half on the face; zero outside)
  and the oracle's readout of the same densities to 1e-10 of the field's
  size;
* B1 rigid translation, B8 uniform strain with a trace and B7 half values
  on a face, each solved on the device: the reference's bars
  (verify_solved_bvp.py:907-935, 1188-1237, 1266-1290), and the readout
  equal to the oracle's readout of the oracle's LU solution to 1e-9 (the
  solutions differ by the FGMRES tolerance).
"""
import numpy as np
import pytest

import _bvp as B

pytestmark = pytest.mark.gpu

NU = (0.25, 0.45)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_somigliana_identity_gpu():
    from mbem_amd import evaluate as ev
    from mbem_amd.cases import box_meshes
    from oracle import evaluate as oe
    model, region, sol, mat = B.exact_density_model(0.45, n_cells=24)
    eps = 0.5
    pts = {"in": B.interior_points(16), "top": B.on_top_points(box_meshes(8)[0])[::2],
           "out": B.outside_points(8)}
    sx = np.linalg.norm(B.exact_sigma(mat))
    for name, q in pts.items():
        s = ev.evaluate_stress(model, region, sol, q, eps, warn_near=False)
        u = ev.evaluate_displacement(model, region, sol, q, eps, warn_near=False)
        so = oe.evaluate_stress(model, region, sol, q, eps)
        uo = oe.evaluate_displacement(model, region, sol, q, eps)
        ue = B.RIGID_C + q @ B.STRAIN.T
        scale = np.linalg.norm(ue, axis=1).max()
        # against the oracle relative to the field's own size (outside the
        # body both readouts are cancellations near zero)
        ds, du = np.abs(s - so).max() / sx, np.abs(u - uo).max() / scale
        print(f"{name}: vs oracle {ds:.1e} / {du:.1e}")
        assert ds < 1e-10 and du < 1e-10
        if name == "in":
            assert B.tensor_error(s, mat) < 1e-3
            assert np.abs(u - ue).max() / scale < 5e-4
        elif name == "top":
            assert abs(B.projections(s, u, q, mat)[0] - 0.5) < 0.03
            assert np.abs(u - 0.5 * ue).max() / scale < 2e-3
        else:
            assert np.abs(s).max() / sx < 2e-3
            assert np.abs(u).max() / scale < 1e-3


def _solve(system, eps):
    from mbem_amd.backends.dense import DenseBackend
    asm = DenseBackend().assemble(system, eps)
    sol = asm.solve(rtol=1e-13, precond_max_dense=0)
    assert asm.report.converged
    return sol


@pytest.mark.parametrize("nu", NU)
def test_b1_rigid_translation_gpu(nu):
    from mbem_amd import evaluate as ev
    from oracle import evaluate as oe
    model, region, system = B.rigid(nu)
    sol = _solve(system, B.EPS)
    mat = region.material
    scale = mat.mu * np.linalg.norm(B.RIGID_C) / 20.0
    tmax = max(np.abs(sol[f"t:{k}"]).max() for k in ("top", "sides", "base")) / scale
    q = B.interior_points()
    u = ev.evaluate_displacement(model, region, sol, q, B.EPS, warn_near=False)
    s = ev.evaluate_stress(model, region, sol, q, B.EPS, warn_near=False)
    print(f"nu={nu}: t {tmax:.1e}, u {np.abs(u - B.RIGID_C).max():.1e}")
    assert tmax < 1e-10
    assert np.abs(u - B.RIGID_C).max() / np.linalg.norm(B.RIGID_C) < 5e-3
    assert np.abs(s).max() / scale < 1e-6
    uo = oe.evaluate_displacement(model, region, B.oracle_solution(system), q, B.EPS)
    assert _rel(u, uo) < 1e-9


@pytest.mark.parametrize("nu", NU)
def test_b8_volumetric_strain_gpu(nu):
    from mbem_amd import evaluate as ev
    from oracle import evaluate as oe
    model, region, system = B.strain(nu)
    mat = region.material
    sol = _solve(system, B.EPS)
    q = B.interior_points()
    s = ev.evaluate_stress(model, region, sol, q, B.EPS, warn_near=False)
    hyd, ten = B.hydrostatic_error(s, mat), B.tensor_error(s, mat)
    so = oe.evaluate_stress(model, region, B.oracle_solution(system), q, B.EPS)
    print(f"nu={nu}: hydrostatic {hyd:.2e}, tensor {ten:.2e}, vs oracle {_rel(s, so):.1e}")
    assert hyd < 3.5e-2 and ten < 3.5e-2
    assert _rel(s, so) < 1e-9


@pytest.mark.parametrize("nu", NU)
@pytest.mark.parametrize("frac", [0.1, 0.025])
def test_b7_half_on_a_face_gpu(nu, frac):
    from mbem_amd import evaluate as ev
    from mbem_amd.kernels.basis import element_sizes
    from oracle import evaluate as oe
    model, region, system = B.half(nu)
    mat = region.material
    top = region.patches[0]
    eps = {"top": frac * element_sizes(top.mesh), "sides": "auto", "base": "auto"}
    sol = _solve(system, eps)
    qc = top.mesh.centroids()
    qc = qc[(np.abs(qc[:, 0]) < 12.0) & (np.abs(qc[:, 1]) < 12.0)]
    qb = B.on_top_points(top.mesh)
    s_c = ev.evaluate_stress(model, region, sol, qc, eps, warn_near=False)
    u_b = ev.evaluate_displacement(model, region, sol, qb, eps, warn_near=False)
    qd = qb.copy()
    qd[:, 2] = -1.5
    s_d = ev.evaluate_stress(model, region, sol, qd, eps, warn_near=False)
    u_d = ev.evaluate_displacement(model, region, sol, qd, eps, warn_near=False)
    s_on = B.projections(s_c, np.zeros_like(qc), qc, mat)[0]
    u_on = B.projections(np.zeros_like(s_d), u_b, qb, mat)[1]
    s_deep, u_deep = B.projections(s_d, u_d, qd, mat)
    print(f"nu={nu} eps/h={frac}: on the face {s_on:.4f} / {u_on:.4f}, deep {s_deep:.4f} / "
          f"{u_deep:.4f}")
    assert abs(s_on - 0.5) < 1.2e-2 and abs(u_on - 0.5) < 1.5e-2
    assert s_deep > 0.97 and u_deep > 0.97
    so = oe.evaluate_stress(model, region, B.oracle_solution(system, eps), qc, eps)
    assert _rel(s_c, so) < 1e-9
