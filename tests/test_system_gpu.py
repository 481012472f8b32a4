"""GPU parity of the dense system, its solves and the readout against the
oracle, on the CPU-runnable box-with-dipping-fault workload (N = 2,048 box
triangles, n = 6,144 unknowns, 400 fault triangles, eps = 0.5 h).

Bars (FP64, 1e-10 relative):
  A, b, calibrated diagonal : every entry within 1e-10 of max|.|, Frobenius
                              within 1e-12;
  LU solution               : within 1e-10 (relative 2-norm) of the oracle's LU;
  LU condition estimate     : within 1e-6 of LAPACK gecon's on the oracle's
                              matrix (the same estimator on other factors of
                              the same matrix);
  preconditioner apply      : z = M r within 1e-10 (relative 2-norm) of the
                              oracle's block Gauss-Seidel sweep (FP64 chunk
                              inverses; 1e-5 with the mixed solve's FP32
                              ones) for cluster chunks of 9,000 and 1,500
                              DOFs and the dense-LU rung;
  FGMRES solution           : converged to true relres 1e-13, then within
                              1e-10 of the oracle's LU, iteration count
                              within 2 of the oracle's MGS FGMRES;
  displacement / stress     : within 1e-10 of the oracle readout on the
                              50 x 50 surface grid and an interior grid.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    from oracle import dense as od
    model, eps, info = box_fault_model(16)
    system = generate_system(model)
    print(f"case: box n_cells=16, n={system.layout.n_unknowns}", flush=True)
    asm = DenseBackend(calibrate_all_rows=True).assemble(system, eps)
    A, b, calib = od.assemble(system, eps)
    x_lu = od.lu_solve(A, b)
    return dict(model=model, eps=eps, system=system, asm=asm, A=A, b=b, calib=calib, x_lu=x_lu)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_assembled_matrix_and_rhs(c1):
    asm = c1["asm"]
    A, b = asm.A, asm.b
    err = np.max(np.abs(A - c1["A"])) / np.max(np.abs(c1["A"]))
    print(f"A: max entry err {err:.2e}, fro {_rel(A, c1['A']):.2e}; "
          f"b: {_rel(b, c1['b']):.2e}; assembly {asm.assembly.seconds:.3f} s", flush=True)
    assert err < 1e-10
    assert _rel(A, c1["A"]) < 1e-12
    assert np.max(np.abs(b - c1["b"])) < 1e-10 * np.max(np.abs(c1["b"]))
    # per slot block, so no block hides under a larger one
    for rs in asm.layout.slots:
        for cs in asm.layout.slots:
            blk = A[rs.offset:rs.stop, cs.offset:cs.stop]
            ref = c1["A"][rs.offset:rs.stop, cs.offset:cs.stop]
            assert np.max(np.abs(blk - ref)) <= 1e-10 * max(np.max(np.abs(ref)), 1e-300)


def test_calibrated_diagonal(c1):
    asm = c1["asm"]
    for key, C in c1["calib"].items():
        G = asm.calib[key]
        assert np.max(np.abs(G - C)) < 1e-10 * np.max(np.abs(C))


def test_lu_solve(c1):
    asm = c1["asm"]
    sol = asm.solve(method="lu")
    x = np.concatenate([sol[s.name].ravel() for s in asm.layout.slots])
    print(f"LU: rel err vs oracle LU {_rel(x, c1['x_lu']):.2e}, residual {asm.report.residual:.2e}")
    assert _rel(x, c1["x_lu"]) < 1e-10
    import scipy.linalg as sl
    from scipy.linalg.lapack import dgecon
    A = c1["A"]
    lu, _ = sl.lu_factor(A)
    rcond, info = dgecon(lu, float(np.abs(A).sum(axis=0).max()), norm="1")
    ref = 1.0 / rcond
    got = asm.report.cond_estimate
    print(f"LU cond estimate {got:.6e}, LAPACK gecon {ref:.6e}", flush=True)
    assert info == 0 and abs(got - ref) <= 1e-6 * ref
    asm.release()


def test_fgmres_block_jacobi(c1):
    from oracle import dense as od
    asm = c1["asm"]
    sol = asm.solve(rtol=1e-13, precond_max_dense=0)
    rep = asm.report
    x = np.concatenate([sol[s.name].ravel() for s in asm.layout.slots])
    print(f"FGMRES: {rep}; rel err vs oracle LU {_rel(x, c1['x_lu']):.2e}", flush=True)
    assert rep.converged and rep.true_relres < 1e-13
    assert _rel(x, c1["x_lu"]) < 1e-10
    pts = {id(p): p.collocation_points() for p in c1["model"].regions[0].patches}
    A = c1["A"]
    M = od.BlockJacobiGS(A, asm.layout, lambda p: pts[id(p)], max_dense=0)
    _, orep = od.fgmres(lambda v: A @ v, c1["b"], M=M, rtol=1e-10)
    asm.release()
    asm.solve(rtol=1e-10, precond_max_dense=0)
    print(f"iterations to 1e-10: GPU {asm.report.iterations}, oracle {orep['iterations']}")
    assert abs(asm.report.iterations - orep["iterations"]) <= 2


@pytest.mark.parametrize("max_dense,bj_chunk,dtype", [(0, 9000, "f64"), (0, 1500, "f64"),
                                                     (4000, 1500, "f64"), (0, 1500, "f32")])
def test_preconditioner_apply_matches_oracle(c1, max_dense, bj_chunk, dtype):
    """la/preconditioner.py:314-543: one forward sweep over the patch
    super-blocks in slot order, each solved by its rung."""
    import torch
    from mbem_amd.la.preconditioner import BlockGaussSeidel
    from oracle import dense as od
    asm = c1["asm"]
    pts = {id(p): p.collocation_points() for p in c1["model"].regions[0].patches}
    M_or = od.BlockJacobiGS(c1["A"], asm.layout, lambda p: pts[id(p)], max_dense=max_dense,
                            bj_chunk=bj_chunk)
    M = BlockGaussSeidel(asm, max_dense=max_dense, bj_chunk=bj_chunk, grouping="patch",
                         inverse_dtype=dtype)
    rng = np.random.default_rng(bj_chunk + max_dense)
    r = rng.standard_normal(asm.layout.n_unknowns)
    z = M(torch.as_tensor(r, device=asm.A_dev.device)).cpu().numpy()
    zr = M_or(r)
    e = _rel(z, zr)
    rungs = sorted({sb.rung for sb in M.sbs})
    print(f"max_dense {max_dense}, chunk {bj_chunk}, {dtype}: rungs {rungs}, "
          f"rel err vs oracle sweep {e:.2e}", flush=True)
    assert e < (1e-10 if dtype == "f64" else 1e-5)


def test_readout(c1):
    from mbem_amd.cases import surface_grid
    from mbem_amd.evaluate import evaluate_displacement, evaluate_stress
    from oracle import evaluate as oe
    from oracle.dense import split
    model = c1["model"]
    region = model.regions[0]
    sol = split(c1["system"].layout, c1["x_lu"])
    eps = c1["eps"]
    grid = surface_grid(50)
    inner = surface_grid(12, half=15.0, z=-4.0 * eps["top"])
    pts = np.vstack([grid, inner])
    u = evaluate_displacement(model, region, sol, pts, eps, warn_near=False)
    ur = oe.evaluate_displacement(model, region, sol, pts, eps)
    print(f"displacement rel err {_rel(u, ur):.2e}")
    assert np.max(np.abs(u - ur)) < 1e-10 * np.max(np.abs(ur))
    s = evaluate_stress(model, region, sol, pts, eps, warn_near=False)
    sr = oe.evaluate_stress(model, region, sol, pts, eps)
    print(f"stress rel err {_rel(s, sr):.2e}")
    assert np.max(np.abs(s - sr)) < 1e-10 * np.max(np.abs(sr))
    s_full = evaluate_stress(model, region, sol, pts, eps, warn_near=False, near_list=False)
    sr_full = oe.evaluate_stress(model, region, sol, pts, eps, near_list=False)
    assert np.max(np.abs(s_full - sr_full)) < 1e-10 * np.max(np.abs(sr_full))


def test_fault_traction(c1):
    """evaluate_fault_traction (evaluate.py:475-492): the stress readout at the
    fault's collocation points dotted with their normals, against the oracle's
    stress at the same points (1e-10 of the largest entry); the fault may be
    named; an unknown name raises."""
    from mbem_amd.evaluate import evaluate_fault_traction
    from oracle import evaluate as oe
    from oracle.dense import split
    model = c1["model"]
    region = model.regions[0]
    fault = region.faults[0]
    sol = split(c1["system"].layout, c1["x_lu"])
    eps = c1["eps"]
    t = evaluate_fault_traction(model, region, sol, fault, eps)
    sr = oe.evaluate_stress(model, region, sol, fault.collocation_points(), eps)
    tr = np.einsum("nij,nj->ni", sr, fault.collocation_normals())
    print(f"fault traction rel err {_rel(t, tr):.2e}")
    assert t.shape == (fault.collocation_points().shape[0], 3)
    assert np.max(np.abs(t - tr)) < 1e-10 * np.max(np.abs(tr))
    assert np.array_equal(evaluate_fault_traction(model, region.name, sol, fault.name, eps), t)
    with pytest.raises(KeyError):
        evaluate_fault_traction(model, region, sol, "no such fault", eps)


def test_assembly_modes(c1):
    """dense.py:1-30, 49-86, 650-687: "basis" and "legacy" assemble the same
    operator as "direct" (here through the same device kernels, so bit for
    bit); legacy keeps the reference's restrictions -- one global scalar eps,
    no material rebuilds -- and an unknown mode or LU flavour is refused."""
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.mesh import ElasticMaterial
    system, eps = c1["system"], c1["eps"]
    direct = DenseBackend().assemble(system, eps)
    basis = DenseBackend(mode="basis").assemble(system, eps)
    assert torch.equal(basis.A_dev, direct.A_dev) and torch.equal(basis.b_dev, direct.b_dev)
    region = c1["model"].regions[0].name
    soft = ElasticMaterial(mu=20.0, lam=30.0)
    assert torch.equal(basis.rebuild_for_materials({region: soft}).A_dev,
                       direct.rebuild_for_materials({region: soft}).A_dev)
    with pytest.raises(ValueError):
        DenseBackend(mode="legacy").assemble(system, eps if isinstance(eps, dict) else {"x": eps})
    h = 0.5
    legacy = DenseBackend(mode="legacy").assemble(system, h)
    assert torch.equal(legacy.A_dev, DenseBackend().assemble(system, h).A_dev)
    with pytest.raises(ValueError):
        legacy.rebuild_for_materials({region: soft})
    with pytest.raises(ValueError):
        DenseBackend(mode="fast")
    with pytest.raises(ValueError):
        DenseBackend(lu="half")


def test_rigid_translation_gives_zero_traction():
    """Manufactured solution (verify_solved_bvp.py B1): prescribe a constant
    displacement on every wall and the base, leave the top free: the solve
    returns u_top = c and zero traction on the walls, to rounding -- the
    calibrated diagonal annihilates constants exactly."""
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    c = np.array([0.3, -0.2, 0.5])
    model, eps, _ = box_fault_model(8, fault=False)
    for p in model.regions[0].patches[1:]:
        p.value = c
    asm = DenseBackend().assemble(generate_system(model), eps)
    sol = asm.solve(method="lu")
    assert np.max(np.abs(sol["u:top"] - c)) < 1e-10
    assert np.max(np.abs(sol["t:sides"])) < 1e-9 * 30.0
    assert np.max(np.abs(sol["t:base"])) < 1e-9 * 30.0


def test_fault_convention_selfcheck():
    from mbem_amd import selfcheck
    selfcheck.ensure_fault_convention("stress")
    rep = selfcheck.report()
    print(rep)
    assert set(rep) == {"core", "stress"}
