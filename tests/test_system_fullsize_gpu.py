"""GPU checks at the benchmark workloads' own sizes, through properties that
need no CPU solve at that size (the oracle's assembly and LU of a 60,000 x
60,000 system would take hours; it pins the same kernels at N = 2,048 in
test_system_gpu.py):

  * N = 20,000 box without fault, constant displacement c on the walls and
    base, free top (the manufactured solution of verify_solved_bvp.py B1):
    u_top = c to 1e-9 and wall tractions below 3e-8 (1e-9 of mu) -- the
    calibrated diagonal annihilates constants exactly;
  * the N = 20,000 box + 676-triangle fault (the headline system): plain FP64
    FGMRES, the mixed solve with int8 and with BF16 far field, and the dense
    LU give the same solution to 1e-9 (relative 2-norm), the iterative ones at
    true residual 1e-12;
  * 10^6 points x 20,676 sources (the readout throughput case): every stress
    finite, and 32 sampled interior points and 32 sampled points within
    3 eps of the fault (where the eigenstress subtraction and the near-field
    terms act) equal to the oracle's C restatement to 1e-10 (relative max);
  * N = 10,368 Laplace states (32 s values x 8 slips, the batched workload):
    all 256 runs converge to 1e-11 and two states equal a direct assembly at
    their moduli + FP64 FGMRES to 1e-8 + 1e-15 lam(s)/mu(s) (the direct
    assembly loses digits like lam/mu; test_laplace_gpu.py explains the bar).
"""
import gc

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _free():
    gc.collect()
    torch.cuda.empty_cache()


def test_rigid_translation_n20000():
    from mbem_amd.backends.dense import DenseBackend
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    c = np.array([0.3, -0.2, 0.5])
    model, eps, info = box_fault_model(50, fault=False)
    print(f"case: box n_cells=50 without fault, N={info['box_triangles']}", flush=True)
    for p in model.regions[0].patches[1:]:
        p.value = c
    asm = DenseBackend().assemble(generate_system(model), eps)
    sol = asm.solve(rtol=1e-13, precond_bj_chunk=1500, precond_grouping="region")
    print(f"{asm.report}", flush=True)
    eu = np.max(np.abs(sol["u:top"] - c))
    et = max(np.max(np.abs(sol["t:sides"])), np.max(np.abs(sol["t:base"])))
    print(f"u_top - c max {eu:.2e}, wall traction max {et:.2e}", flush=True)
    assert eu < 1e-9 and et < 1e-9 * 30.0
    asm.release()
    del asm
    _free()


def test_headline_system_all_solvers_agree():
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.la.mixed import solve_mixed
    from mbem_amd.model import generate_system
    model, eps, info = box_fault_model(50, fault_cells=(13, 26))
    system = generate_system(model)
    print(f"case: headline box + fault, N={info['box_triangles']}, "
          f"n={system.layout.n_unknowns}", flush=True)
    asm = AssembledDense(system, eps, "direct", jump="calibrated")
    x_p = asm.solve(rtol=1e-12, precond_bj_chunk=1500, precond_grouping="region",
                    return_device=True).cpu().numpy()
    rep_p = asm.report
    print(f"plain: {rep_p}", flush=True)
    assert rep_p.converged and rep_p.true_relres < 1e-12
    out = {}
    for storage in ("int8", "bf16"):
        x, rep = solve_mixed(asm, rtol=1e-12, storage=storage)
        out[storage] = x.cpu().numpy()
        print(f"mixed {storage}: {rep}; vs plain {_rel(out[storage], x_p):.2e}", flush=True)
        assert rep.converged and rep.outer.true_relres < 1e-12
        del x
        _free()
    asm.release()
    x_lu = asm.solve(method="lu", return_device=True).cpu().numpy()
    print(f"lu: {asm.report}; plain vs lu {_rel(x_p, x_lu):.2e}", flush=True)
    assert _rel(x_p, x_lu) < 1e-9
    for storage, x in out.items():
        assert _rel(x, x_lu) < 1e-9, storage
    asm.release()
    del asm
    _free()


def test_readout_throughput_case_symmetry_and_sample():
    from mbem_amd.cases import readout_case
    from mbem_amd.kernels import device as dv
    from oracle import tri_kernels as ok
    mu, lam = 30.0, 30.0
    pts, tv, nrm, eps, dens, info = readout_case(1_000_000)
    print(f"case: {pts.shape[0]} points x {tv.shape[0]} sources", flush=True)
    dev = torch.device("cuda")
    src = dv.Sources(torch.from_numpy(tv).to(dev), torch.from_numpy(nrm).to(dev),
                     torch.from_numpy(eps).to(dev))
    X = torch.from_numpy(pts).to(dev)
    D = torch.from_numpy(dens).to(dev)
    out = torch.empty((X.shape[0], 3, 3), dtype=torch.float64, device=dev)
    chunk = 131072
    for p0 in range(0, X.shape[0], chunk):
        p1 = min(p0 + chunk, X.shape[0])
        dv.stress("dd", X[p0:p1], src, D, mu, lam, out[p0:p1])
        dv.stress("eigen", X[p0:p1], src, D, mu, lam, out[p0:p1], scale=-1.0, beta=1.0)
    S = out.cpu().numpy()
    assert np.all(np.isfinite(S))
    # readout_case puts the interior points first and the near-fault half after
    half = pts.shape[0] // 2
    rng = np.random.default_rng(17)
    sel = np.concatenate([rng.choice(half, size=32, replace=False),
                          half + rng.choice(pts.shape[0] - half, size=32, replace=False)])
    ref = ok.dd_stress_contract(pts[sel], tv, nrm, eps, dens, mu, lam) \
        - ok.eigenstress_contract(pts[sel], tv, nrm, eps, dens, mu, lam)
    err = float(np.abs(S[sel] - ref).max() / np.abs(ref).max())
    print(f"64 sampled points (32 near the fault) vs oracle: max rel err {err:.2e}", flush=True)
    assert err < 1e-10
    del out, X, D, src
    _free()


def test_laplace_states_full_size():
    from mbem_amd import laplace, rheology as rh
    from mbem_amd.backends.dense import AssembledDense
    from mbem_amd.cases import box_fault_model
    from mbem_amd.kernels import basis as kb
    from mbem_amd.la.solver import fgmres
    from mbem_amd.model import generate_system
    from mbem_amd.viscoelastic import LaplaceStates
    model, _, info = box_fault_model(36, fault_cells=(9, 18))
    region = model.regions[0]
    h = float(np.mean([kb.element_sizes(p.mesh).mean() for p in region.patches]))
    eps = {p.name: 0.5 * h for p in region.patches}
    eps["fault"] = 0.5 * float(kb.element_sizes(region.faults[0].mesh).mean())
    system = generate_system(model)
    tau = 1.0e19 / 3.0e10
    grid = laplace.s_grid(0.01 * tau, 2e7 * tau)
    states = rh.laplace_states({"crust": (30.0, 30.0)}, {"crust": rh.Maxwell(tau)}, grid)[1:]
    rng = np.random.default_rng(5)
    nf = region.faults[0].mesh.n_triangles
    slips = []
    for _ in range(8):
        v = rng.normal(size=(nf, 3))
        slips.append({"fault": v / np.linalg.norm(v, axis=1, keepdims=True)})
    print(f"case: Laplace states N={info['box_triangles']} n={system.layout.n_unknowns}, "
          f"{len(states)} states x {len(slips)} slips", flush=True)
    ls = LaplaceStates(system, eps, {"crust": rh.Maxwell(tau)}, bj_chunk=1500)
    x, rep = ls.solve(states, slips, rtol=1e-11)
    print(f"{rep}", flush=True)
    assert rep.batched.converged.all()
    b0, b1 = ls.rhs_parts(slips)
    for i in (0, len(states) - 1):
        st = states[i]
        mat = st.materials.get("crust")
        asm = AssembledDense(system, eps, _materials={"crust": mat} if mat is not None else None)
        mu, lam, g = ls.state_moduli(st.s)
        tol = 1e-8 + 1e-15 * lam / mu
        # the first random slip's right-hand side at this state
        b = (b0 + g * b1)[:, 0].contiguous()
        M = asm.preconditioner(0, 1500, grouping="region")
        x_d, r = fgmres(asm, b, M=M, rtol=1e-13, restart=200, maxiter=600)
        e0 = _rel(x[i, 0], x_d.cpu().numpy())
        print(f"state {st.label} s*tau={st.s * tau:.3g}: vs direct assembly + FGMRES "
              f"{e0:.2e} (bar {tol:.1e}; {r})", flush=True)
        assert r.converged and e0 < tol
        asm.release()
        del asm, M
        _free()
    ls.release()
    del ls
    _free()
